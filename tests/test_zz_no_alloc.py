"""GPU checks of hbs_gemm_fp8res_noalloc, the allocation-free form of
hbs_gemm_fp8res (its workspace lives in a fixed scratch array inside the
library).  Called through the C-ABI; every result must equal the workspace
entry point's bit for bit, since both run the same three kernels on the same
workspace layout.  The scratch keeps the previous call's contents, so the
second call of each pair is a smaller problem than the first."""
import ctypes

import pytest

torch = pytest.importorskip("torch")
import hb_hammerbench_amd as hb  # noqa: E402

pytestmark = pytest.mark.gpu


def dev_fill(n, kind, seed):
    x = torch.empty(int(n), dtype=torch.float32, device="cuda")
    hb.fill_random(x, n, kind, seed)
    return x


def both(M, N, K, lda, ldb, ldc, seed, kind=hb.GEN_NORMAL_OUTLIER):
    """C from the workspace entry point and from the allocation-free one,
    both called through the C-ABI on the current stream."""
    a = dev_fill(M * lda, kind, seed)
    b = dev_fill(K * ldb, kind, seed + 1)
    c_ws = torch.full((M * ldc,), -7.0, dtype=torch.float32, device="cuda")
    c_na = torch.full((M * ldc,), -7.0, dtype=torch.float32, device="cuda")
    L = hb.lib()
    s = torch.cuda.current_stream().cuda_stream
    need = L.hbs_gemm_fp8res_workspace_size(M, N, K)
    assert 0 < need <= hb.FP8RES_NOALLOC_SCRATCH_BYTES
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    rc = L.hbs_gemm_fp8res(a.data_ptr(), b.data_ptr(), c_ws.data_ptr(), M, N, K, lda, ldb, ldc,
                           ws.data_ptr(), need, s)
    assert rc == 0, hb.lib().hbs_strerror(rc)
    rc = L.hbs_gemm_fp8res_noalloc(a.data_ptr(), b.data_ptr(), c_na.data_ptr(), M, N, K, lda,
                                   ldb, ldc, s)
    assert rc == 0, hb.lib().hbs_strerror(rc)
    torch.cuda.synchronize()
    return c_ws, c_na


@pytest.mark.parametrize("first,second", [
    ((1000, 777, 1500, 1500, 777, 777), (300, 260, 200, 200, 260, 260)),
    # strided operands, then a smaller contiguous problem
    ((513, 640, 1028, 1040, 644, 648), (64, 64, 64, 64, 64, 64)),
])
def test_noalloc_equals_workspace_entry_point_twice(first, second):
    for i, shape in enumerate((first, second)):
        M, N, K, lda, ldb, ldc = shape
        print(f"\nnoalloc call {i}: M={M} N={N} K={K} lda={lda} ldb={ldb} ldc={ldc}", flush=True)
        c_ws, c_na = both(M, N, K, lda, ldb, ldc, seed=100 + i)
        assert torch.equal(c_ws, c_na), shape
        # the padding between rows (ldc > N) is left untouched by both
        if ldc > N:
            assert bool((c_na.view(M, ldc)[:, N:] == -7.0).all())


def test_noalloc_near_the_scratch_limit_then_small():
    """M = N = 4096, K = 2560 needs 60 MiB of the 64 MiB scratch."""
    for i, (M, N, K) in enumerate([(4096, 4096, 2560), (384, 256, 130)]):
        print(f"\nnoalloc call {i}: M={M} N={N} K={K}", flush=True)
        c_ws, c_na = both(M, N, K, K, N, N, seed=200 + i)
        assert torch.equal(c_ws, c_na), (M, N, K)


def test_noalloc_pattern_bit_exact():
    """The reference's integer pattern (apps/sgemm/main.cpp:46-52) stays exact
    through power-of-two tile scales, as for hbs_gemm_fp8res."""
    import numpy as np

    import oracle
    N = 512
    a = torch.empty(N * N, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    c = torch.empty_like(a)
    hb.fill_pattern(a, b, N * N)
    hb.gemm_fp8res_noalloc(a, b, c, N, N, N)
    torch.cuda.synchronize()
    assert np.array_equal(c.cpu().numpy().reshape(N, N), oracle.pattern_slice(N, 0))


def test_noalloc_rejects_problems_beyond_the_scratch():
    P = ctypes.c_void_p(16)
    L = hb.lib()
    assert L.hbs_gemm_fp8res_workspace_size(4096, 4096, 4096) > hb.FP8RES_NOALLOC_SCRATCH_BYTES
    assert L.hbs_gemm_fp8res_noalloc(P, P, P, 4096, 4096, 4096, 4096, 4096, 4096, None) == -7
    with pytest.raises(hb.HbsError) as ei:
        hb.gemm_fp8res_noalloc(16, 16, 16, 8192, 8192, 2048)
    assert ei.value.code == -7
