"""GPU parity of the stream-K GEMM (hbs_gemm_streamk), through the C-ABI.

No reference function exists for stream-K scheduling; it computes the same
product as hbs_sgemm_strided / hbs_gemm_bf16x3 with tiles split over several
workgroups along K.  Pinned by
 * bit-exactness on the reference's integer pattern (apps/sgemm/main.cpp:46-52:
   every partial sum is an integer below 2^24, so no order of additions can
   round);
 * per-element FP64 bounds with the split's extra additions counted: a tile
   split into s segments adds s partial sums, so FP32 |C - R| <= gamma_{K+s+2} S
   and BF16x3 |C - R| <= (3.1 * 2^-16 + gamma_{3K+s+2}) S, S = |A||B|
   (s <= G + 1 for the G workgroups the workspace size reveals);
 * determinism (the last arriver sums in K order), counters left at zero, a
   workspace reused across shapes, and the argument checks.
"""
import numpy as np
import pytest

import oracle

torch = pytest.importorskip("torch")
import hb_hammerbench_amd as hb  # noqa: E402

pytestmark = pytest.mark.gpu
U = 2.0 ** -24
SENTINEL = -9.0


def gamma(n):
    return n * U / (1 - n * U)


def max_segments(M, N, K, prec):
    """G + 1: the workspace is 256 KiB of counters and two 256 KiB slots per
    workgroup, and a tile meets at most every workgroup's run plus one."""
    ws = hb.streamk_workspace_size(M, N, K, prec)
    return (ws - 256 * 1024) // (2 * 256 * 256 * 4) + 1


def bound(prec, M, N, K, S):
    s = max_segments(M, N, K, prec)
    if prec == hb.PREC_FP32:
        return gamma(K + s + 2) * S
    return (3.1 * 2.0 ** -16 + gamma(3 * K + s + 2)) * S


def operands(M, N, K, lda, ldb, ldc, seed, offs=(0, 0, 0)):
    oa, ob, oc = offs
    a = torch.empty(oa + M * lda, dtype=torch.float32, device="cuda")
    b = torch.empty(ob + K * ldb, dtype=torch.float32, device="cuda")
    c = torch.full((oc + M * ldc,), SENTINEL, dtype=torch.float32, device="cuda")
    hb.fill_random(a, a.numel(), hb.GEN_UNIFORM, seed)
    hb.fill_random(b, b.numel(), hb.GEN_LOGUNIFORM, seed + 1)
    return a, b, c


def views(a, b, c, M, N, K, lda, ldb, ldc, offs=(0, 0, 0)):
    oa, ob, oc = offs
    Ah, Bh, Ch = a.cpu().numpy(), b.cpu().numpy(), c.cpu().numpy()
    A = np.lib.stride_tricks.as_strided(Ah[oa:], (M, K), (lda * 4, 4)).copy()
    B = np.lib.stride_tricks.as_strided(Bh[ob:], (K, N), (ldb * 4, 4)).copy()
    C = np.lib.stride_tricks.as_strided(Ch[oc:], (M, N), (ldc * 4, 4)).copy()
    return A, B, C, Ch


# (M, N, K, pads, pointer offsets): 36 tiles over 216 (FP32) or 144 (BF16x3)
# workgroups (each tile split six or four ways); ragged M and N (clamped loads); K not a multiple of 32
# (zero fill past K in the last segment); unaligned operands (predicated
# loads); runs that do not divide a tile's stages; a deep K over few tiles
# (every CU, tiles split 16 ways or more); fewer units than CUs.
CASES = [
    (1536, 1536, 1536, (0, 0, 0), (0, 0, 0)),
    (1000, 1300, 1024, (4, 4, 0), (0, 0, 0)),
    (700, 900, 1002, (2, 0, 6), (0, 0, 0)),
    (513, 257, 777, (1, 3, 2), (1, 2, 3)),
    (2048, 768, 2048, (0, 0, 0), (0, 0, 0)),
    (256, 512, 40000, (0, 4, 0), (0, 0, 0)),
    (300, 200, 96, (0, 0, 0), (0, 0, 0)),
    (5, 7, 3, (0, 1, 0), (0, 0, 0)),
]


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
@pytest.mark.parametrize("case", CASES)
def test_streamk_within_bound_and_padding_untouched(prec, case):
    M, N, K, (pa, pb, pc), offs = case
    lda, ldb, ldc = K + pa, N + pb, N + pc
    print(f"streamk prec {prec} case {case}", flush=True)
    a, b, c = operands(M, N, K, lda, ldb, ldc, 7, offs)
    oa, ob, oc = offs
    hb.gemm_streamk(a[oa:], b[ob:], c[oc:], M, N, K, prec, lda, ldb, ldc)
    torch.cuda.synchronize()
    A, B, C, Ch = views(a, b, c, M, N, K, lda, ldb, ldc, offs)
    R, S = oracle.gemm_f64(A, B)
    err = np.abs(C.astype(np.float64) - R)
    assert np.all(err <= bound(prec, M, N, K, S)), float(err.max())
    written = np.zeros(Ch.size, bool)
    for r in range(M):
        written[oc + r * ldc: oc + r * ldc + N] = True
    assert np.all(Ch[~written] == SENTINEL)


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
def test_streamk_reference_pattern_bit_exact(prec):
    """N = 1536 on the reference's i%7 / i%3 inputs, tiles split six or four ways:
    exact, as in the one-tile-per-workgroup kernels."""
    N = 1536
    a = torch.empty(N * N, device="cuda"); b = torch.empty(N * N, device="cuda")
    hb.fill_pattern(a, b, N * N)
    c = torch.empty(N * N, device="cuda")
    hb.gemm_streamk(a, b, c, N, N, N, prec)
    torch.cuda.synchronize()
    A = a.cpu().numpy().reshape(N, N).astype(np.float64)
    B = b.cpu().numpy().reshape(N, N).astype(np.float64)
    assert np.array_equal(c.cpu().numpy().reshape(N, N), (A @ B).astype(np.float32))


def test_streamk_deterministic_counters_zero_and_workspace_reuse():
    """Two calls give the same bits; the counter region (the first 256 KiB of
    the workspace) is zero after each call; the same workspace then serves a
    smaller product of another shape."""
    M = N = K = 1536
    a, b, c = operands(M, N, K, K, N, N, 21)
    ws = hb.gemm_streamk(a, b, c, M, N, K, hb.PREC_BF16X3)
    torch.cuda.synchronize()
    first = c.clone()
    assert not ws[:256 * 1024].any().item()
    hb.gemm_streamk(a, b, c, M, N, K, hb.PREC_BF16X3, workspace=ws)
    torch.cuda.synchronize()
    assert torch.equal(c.view(torch.int32), first.view(torch.int32))
    assert not ws[:256 * 1024].any().item()
    m, n, k = 700, 900, 1002
    a2, b2, c2 = operands(m, n, k, k, n, n, 23)
    hb.gemm_streamk(a2, b2, c2, m, n, k, hb.PREC_FP32, workspace=ws)
    torch.cuda.synchronize()
    A, B, C, _ = views(a2, b2, c2, m, n, k, k, n, n)
    R, S = oracle.gemm_f64(A, B)
    assert np.all(np.abs(C.astype(np.float64) - R) <= bound(hb.PREC_FP32, m, n, k, S))
    assert not ws[:256 * 1024].any().item()


def test_streamk_argument_checks():
    L = hb.lib()
    slot = 256 * 256 * 4
    # 36 tiles: FP32 splits each over 6 workgroups, BF16x3 (whose stages are
    # shorter against the same fix-up) over 4
    need = hb.streamk_workspace_size(1536, 1536, 1536)
    assert need == 256 * 1024 + 2 * 216 * slot
    assert hb.streamk_workspace_size(1536, 1536, 1536, hb.PREC_BF16X3) == 256 * 1024 + 2 * 144 * slot
    for prec in (hb.PREC_FP32, hb.PREC_BF16X3):
        assert hb.streamk_workspace_size(4096, 4096, 4096, prec) == 256 * 1024 + 2 * 256 * slot
        assert hb.streamk_workspace_size(300, 200, 96, prec) == 256 * 1024 + 2 * 2 * slot
        assert hb.streamk_workspace_size(0, 8, 8, prec) == 0
        assert hb.streamk_workspace_size(300 * 256, 300 * 256, 32, prec) == 0  # 90000 tiles
    assert hb.streamk_workspace_size(64, 64, 64, 2) == 0
    a = torch.zeros(64 * 64, device="cuda")
    ws = torch.zeros(need, dtype=torch.uint8, device="cuda")
    ptr = a.data_ptr()
    assert L.hbs_gemm_streamk(2, ptr, ptr, ptr, 64, 64, 64, 64, 64, 64, ws.data_ptr(), need,
                              None) == -1  # precision
    small = hb.streamk_workspace_size(64, 64, 64) - 1
    assert L.hbs_gemm_streamk(0, ptr, ptr, ptr, 64, 64, 64, 64, 64, 64, ws.data_ptr(), small,
                              None) == -4  # HBS_ERR_WORKSPACE
    assert L.hbs_gemm_streamk(0, ptr, ptr, ptr, 64, 64, 64, 64, 64, 64, ws.data_ptr() + 4,
                              need - 4, None) == -3  # HBS_ERR_ALIGN
    assert L.hbs_gemm_streamk(0, ptr, ptr, ptr, 64, 64, 64, 32, 64, 64, ws.data_ptr(), need,
                              None) == -1  # lda < K


@pytest.mark.parametrize("prec", [hb.PREC_FP32, hb.PREC_BF16X3])
def test_streamk_graph_replays_match_eager(prec):
    """Captured into a graph (the counters are reset by the kernel itself, so
    the graph holds the one launch and no memset), each of four replays, with
    the program's own work and synchronisations between them, gives the eager
    call's bits and leaves the counters zero."""
    M, N, K = 1000, 1300, 1024
    a, b, c = operands(M, N, K, K, N, N, 31)
    ws = torch.zeros(hb.streamk_workspace_size(M, N, K, prec), dtype=torch.uint8, device="cuda")
    hb.gemm_streamk(a, b, c, M, N, K, prec, workspace=ws)
    torch.cuda.synchronize()
    eager = c.clone()
    side = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        hb.gemm_streamk(a, b, c, M, N, K, prec, workspace=ws)
    for rep in range(4):
        c.fill_(SENTINEL)
        junk = torch.randn(1 << 20, device="cuda").sum()  # other work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(c.view(torch.int32), eager.view(torch.int32)), rep
        assert not ws[:256 * 1024].any().item(), rep
        del junk
