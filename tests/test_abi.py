"""C-ABI contract checks that need no GPU: the library loads, exports every
symbol include/hbsgemm.h declares, and rejects invalid arguments with the
documented status codes before touching a device (the counterpart of
BSG_CUDA_CALL returning early, apps/sgemm/main.cpp:61-143)."""
import ctypes
import os
import re

import numpy as np

import pytest

import hb_hammerbench_amd as hb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hbsgemm.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hbs_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    syms = declared_symbols()
    for s in ["hbs_sgemm", "hbs_sgemm_strided", "hbs_fill_pattern", "hbs_fill_random",
              "hbs_verify_f64", "hbs_strerror", "hbs_version"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(hb.library_path())
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_python_binding_covers_every_declared_symbol():
    assert set(declared_symbols()) <= set(hb.SIGNATURES)


def test_strerror_and_version():
    L = hb.lib()
    assert L.hbs_version() == 500  # 0.5.0: FP8 workspace pads B to 256 columns; grouped and stream-K BF16x3 on 16x16 MFMAs
    for code in [0, -1, -2, -3, -4, -5, -6, -7, -8]:
        assert L.hbs_strerror(code)
    assert L.hbs_strerror(-99) == b"unknown status code"


P = ctypes.c_void_p(16)  # non-null dummy pointer; never dereferenced on these paths


@pytest.mark.parametrize("N,niter", [(0, 1), (-3, 1), (16, 0), (16, -1)])
def test_sgemm_rejects_bad_sizes(N, niter):
    assert hb.lib().hbs_sgemm(P, P, P, N, niter, None) == -1


def test_sgemm_rejects_null_pointers():
    L = hb.lib()
    assert L.hbs_sgemm(None, P, P, 16, 1, None) == -2
    assert L.hbs_sgemm(P, None, P, 16, 1, None) == -2
    assert L.hbs_sgemm(P, P, None, 16, 1, None) == -2


def test_sgemm_rejects_overflowing_batch():
    assert hb.lib().hbs_sgemm(P, P, P, 1 << 30, 1 << 30, None) == -1


def test_strided_rejects_bad_leading_dimensions():
    L = hb.lib()
    assert L.hbs_sgemm_strided(P, P, P, 8, 8, 8, 7, 8, 8, 0, 0, 0, 1, None) == -1
    assert L.hbs_sgemm_strided(P, P, P, 8, 8, 8, 8, 7, 8, 0, 0, 0, 1, None) == -1
    assert L.hbs_sgemm_strided(P, P, P, 8, 8, 8, 8, 8, 7, 0, 0, 0, 1, None) == -1
    assert L.hbs_sgemm_strided(P, P, P, 8, 8, 8, 8, 8, 8, -1, 0, 0, 1, None) == -1
    assert L.hbs_sgemm_strided(P, P, P, 8, 8, 8, 8, 8, 8, 0, 0, 0, 0, None) == -1


def test_fill_and_verify_reject_bad_arguments():
    L = hb.lib()
    assert L.hbs_fill_pattern(P, P, -1, 0, None) == -1
    assert L.hbs_fill_pattern(P, P, 0, 0, None) == 0  # empty fill is a no-op
    assert L.hbs_fill_random(P, 10, 7, 1, 0, None) == -1
    assert L.hbs_fill_random(None, 10, 0, 1, 0, None) == -2
    out = hb.ErrStats()
    assert L.hbs_verify_f64(P, P, P, 8, 8, 8, 8, 8, 8, 0, 9, ctypes.byref(out), None) == -1
    assert L.hbs_verify_f64(P, P, P, 8, 8, 8, 8, 8, 8, 4, 4, ctypes.byref(out), None) == -1
    assert L.hbs_verify_f64(P, P, P, 8, 8, 8, 8, 8, 8, 0, 8, None, None) == -2


def test_python_wrapper_raises_hbs_error():
    with pytest.raises(hb.HbsError) as ei:
        hb.sgemm(P.value, P.value, P.value, 0, 1, stream=0)
    assert ei.value.code == -1


def test_bf16x3_argument_checks():
    """The BF16x3 entries take no workspace; argument errors come before any
    device call."""
    L = hb.lib()
    assert L.hbs_sgemm_bf16x3(P, P, P, 0, 1, None) == -1
    assert L.hbs_sgemm_bf16x3(P, P, P, 64, 0, None) == -1
    assert L.hbs_sgemm_bf16x3(P, P, P, 1 << 30, 1 << 20, None) == -1  # N * N * niter overflows
    assert L.hbs_sgemm_bf16x3(None, P, P, 64, 1, None) == -2
    assert L.hbs_sgemm_bf16x3(P, P, None, 64, 1, None) == -2
    assert L.hbs_gemm_bf16x3(P, P, P, 0, 8, 8, 8, 8, 8, None) == -1
    assert L.hbs_gemm_bf16x3(P, P, P, 8, 8, 8, 7, 8, 8, None) == -1  # lda < K
    assert L.hbs_gemm_bf16x3(P, P, P, 8, 8, 8, 8, 7, 8, None) == -1  # ldb < N
    assert L.hbs_gemm_bf16x3(P, P, P, 8, 8, 8, 8, 8, 7, None) == -1  # ldc < N
    assert L.hbs_gemm_bf16x3(P, None, P, 8, 8, 8, 8, 8, 8, None) == -2


def test_grouped_plan_argument_checks():
    L = hb.lib()
    h = ctypes.c_void_p()
    good = hb.Problem(16, 32, 48, 8, 8, 8, 8, 8, 8)
    arr = (hb.Problem * 1)(good)
    assert L.hbs_grouped_plan_create(arr, 0, 0, None, ctypes.byref(h)) == -1
    assert L.hbs_grouped_plan_create(None, 1, 0, None, ctypes.byref(h)) == -2
    assert L.hbs_grouped_plan_create(arr, 1, 7, None, ctypes.byref(h)) == -1
    assert L.hbs_grouped_plan_create(arr, 1, 0, None, None) == -2
    # any shape and alignment passes validation (the dry run needs no device)
    info = hb.GroupedInfo()
    for odd in (hb.Problem(16, 32, 48, 8, 6, 8, 8, 6, 8),      # N % 4
                hb.Problem(20, 32, 48, 8, 8, 8, 8, 8, 8)):     # A misaligned
        arr1 = (hb.Problem * 1)(odd)
        assert L.hbs_grouped_plan_describe(arr1, 1, 0, None, -1, 1, 256, ctypes.byref(info)) == 0
        assert info.tiles == 1 and info.workgroups == 1
    bad = (hb.Problem * 1)(hb.Problem(16, 32, 48, 8, 8, 8, 4, 8, 8))    # lda < K
    assert L.hbs_grouped_plan_create(bad, 1, 0, None, ctypes.byref(h)) == -1
    bad = (hb.Problem * 1)(hb.Problem(None, 32, 48, 8, 8, 8, 8, 8, 8))
    assert L.hbs_grouped_plan_create(bad, 1, 0, None, ctypes.byref(h)) == -2
    assert L.hbs_grouped_plan_create(arr, 1, 0, b"/nonexistent/tuning.txt", ctypes.byref(h)) == -1
    assert L.hbs_grouped_plan_create_forced(arr, 1, 0, 10, 1, ctypes.byref(h)) == -1  # ids 0..9
    assert L.hbs_grouped_plan_create_forced(arr, 1, 1, 5, 1, ctypes.byref(h)) == -1  # FP32 only
    assert L.hbs_grouped_plan_create_forced(arr, 1, 1, 6, 1, ctypes.byref(h)) == -1  # FP32 only
    assert L.hbs_grouped_plan_create_forced(arr, 1, 0, 0, 0, ctypes.byref(h)) == -1
    assert L.hbs_grouped_run(None, None) == -2
    assert L.hbs_grouped_plan_destroy(None) == 0
    assert L.hbs_grouped_plan_info(None, None) == -2
    assert L.hbs_grouped_graph_create(None, ctypes.byref(h)) == -2
    assert L.hbs_graph_launch(None, None) == -2
    assert L.hbs_graph_destroy(None) == 0


def test_fp8res_argument_checks():
    L = hb.lib()
    assert L.hbs_gemm_fp8res_workspace_size(0, 8, 8) == 0
    need = L.hbs_gemm_fp8res_workspace_size(16384, 16384, 4096)
    assert need >= 2 * 16384 * 4096 * 3  # e4m3 byte + bf16 residual per element, A and B
    assert L.hbs_gemm_fp8res(P, P, P, 0, 8, 8, 8, 8, 8, P, 1 << 30, None) == -1
    assert L.hbs_gemm_fp8res(P, P, P, 8, 8, 8, 7, 8, 8, P, 1 << 30, None) == -1
    assert L.hbs_gemm_fp8res(None, P, P, 8, 8, 8, 8, 8, 8, P, 1 << 30, None) == -2
    assert L.hbs_gemm_fp8res(P, P, P, 8, 8, 8, 8, 8, 8, None, 1 << 30, None) == -4
    assert L.hbs_gemm_fp8res(P, P, P, 8, 8, 8, 8, 8, 8, P, 16, None) == -4
    assert L.hbs_gemm_fp8res(P, P, P, 8, 8, 8, 8, 8, 8, ctypes.c_void_p(8), 1 << 30, None) == -3


def test_fp8res_workspace_follows_documented_layout():
    """The workspace size is the layout include/hbsgemm.h documents: Mp and Np
    rounded up to 256, Kp to 128, every plane starting on a 256-byte boundary
    (B is padded to the wide kernel's 256 columns whichever form runs)."""
    L = hb.lib()

    def rup(x, m):
        return (x + m - 1) // m * m
    for M, N, K in [(1, 1, 1), (300, 200, 100), (512, 384, 1024), (16384, 640, 4096),
                    (16384, 16384, 4096), (257, 129, 129)]:
        Mp, Np, Kp = rup(M, 256), rup(N, 256), rup(K, 128)
        end = 0
        for nbytes in (Mp * Kp, Mp * Kp * 2, (Mp // 128) * (Kp // 128) * 4,
                       Np * Kp, Np * Kp * 2, (Np // 128) * (Kp // 128) * 4):
            end = rup(end + nbytes, 256)
        assert L.hbs_gemm_fp8res_workspace_size(M, N, K) == end, (M, N, K)


def test_fp8res_noalloc_argument_checks():
    """Argument and size errors of the allocation-free entry point come before
    any device call: problems whose workspace exceeds the fixed scratch get
    HBS_ERR_SCRATCH (-7)."""
    L = hb.lib()
    cap = hb.FP8RES_NOALLOC_SCRATCH_BYTES
    assert L.hbs_gemm_fp8res_workspace_size(4096, 4096, 2688) <= cap
    assert L.hbs_gemm_fp8res_workspace_size(4096, 4096, 2816) > cap
    assert L.hbs_gemm_fp8res_noalloc(P, P, P, 4096, 4096, 2816, 2816, 4096, 4096, None) == -7
    assert L.hbs_gemm_fp8res_noalloc(P, P, P, 0, 8, 8, 8, 8, 8, None) == -1
    assert L.hbs_gemm_fp8res_noalloc(P, P, P, 8, 8, 8, 7, 8, 8, None) == -1
    assert L.hbs_gemm_fp8res_noalloc(P, None, P, 8, 8, 8, 8, 8, 8, None) == -2
    assert b"scratch" in L.hbs_strerror(-7)


def test_square_and_rectangular_entries_reject_sizes_beyond_the_int_limit():
    """M, N, K or a leading dimension above HBS_MAX_DIM would overflow the
    kernels' 32-bit tile arithmetic: HBS_ERR_ARG before any device call."""
    L = hb.lib()
    big = 0x7fffffff - 4096 + 1
    assert L.hbs_sgemm(P, P, P, big, 1, None) == -1
    assert L.hbs_sgemm_strided(P, P, P, big, 8, 8, 8, 8, 8, 0, 0, 0, 1, None) == -1
    assert L.hbs_sgemm_strided(P, P, P, 8, 8, 8, 8, 8, big, 0, 0, 0, 1, None) == -1
    assert L.hbs_sgemm_bf16x3(P, P, P, big, 1, None) == -1
    assert L.hbs_gemm_bf16x3(P, P, P, 8, 8, big, big, 8, 8, None) == -1
    assert L.hbs_gemm_fp8res(P, P, P, 8, big, 8, 8, big, big, P, 1 << 30, None) == -1
    assert L.hbs_gemm_fp8res_workspace_size(8, 8, big) == 0
    out = hb.ErrStats()
    assert L.hbs_verify_f64(P, P, P, big, 8, 8, 8, 8, 8, 0, 8, ctypes.byref(out), None) == -1


def test_streamk_plan_and_argument_checks():
    """The stream-K workspace follows its plan (256 KiB of counters, then two
    256 KiB partial slots for each of G workgroups, 1 <= G <= min(CUs, units));
    argument, size and workspace errors come before any device call."""
    L = hb.lib()
    slot, counters = 256 * 256 * 4, 256 * 1024
    cus = 256  # MI355X, and the planner's default when no device answers
    rng = np.random.default_rng(5)
    for prec in (hb.PREC_FP32, hb.PREC_BF16X3):
        for _ in range(40):
            M, N, K = (int(x) for x in rng.integers(1, 9000, 3))
            ws = L.hbs_gemm_streamk_workspace_size(prec, M, N, K)
            units = -(-M // 256) * -(-N // 256) * -(-K // 32)
            G, rem = divmod(ws - counters, 2 * slot)
            assert rem == 0 and 1 <= G <= min(cus, units), (M, N, K, ws)
        # products that fill the GPU take one workgroup per CU
        assert L.hbs_gemm_streamk_workspace_size(prec, 4096, 4096, 4096) == counters + 2 * cus * slot
        # one stage of one tile: one workgroup
        assert L.hbs_gemm_streamk_workspace_size(prec, 5, 7, 3) == counters + 2 * slot
        assert L.hbs_gemm_streamk_workspace_size(prec, 0, 8, 8) == 0
        assert L.hbs_gemm_streamk_workspace_size(prec, 8, 8, 0x7fffffff) == 0
        assert L.hbs_gemm_streamk_workspace_size(prec, 300 * 256, 300 * 256, 32) == 0  # > 65536 tiles
    assert L.hbs_gemm_streamk_workspace_size(2, 64, 64, 64) == 0
    need = L.hbs_gemm_streamk_workspace_size(0, 64, 64, 64)
    W = ctypes.c_void_p(1 << 20)  # never dereferenced: every call below fails its checks
    assert L.hbs_gemm_streamk(2, P, P, P, 64, 64, 64, 64, 64, 64, W, need, None) == -1
    assert L.hbs_gemm_streamk(0, P, P, P, 0, 64, 64, 64, 64, 64, W, need, None) == -1
    assert L.hbs_gemm_streamk(0, P, P, P, 64, 64, 64, 32, 64, 64, W, need, None) == -1
    assert L.hbs_gemm_streamk(0, None, P, P, 64, 64, 64, 64, 64, 64, W, need, None) == -2
    assert L.hbs_gemm_streamk(0, P, P, P, 64, 64, 64, 64, 64, 64, W, need - 1, None) == -4
    assert L.hbs_gemm_streamk(0, P, P, P, 64, 64, 64, 64, 64, 64, None, need, None) == -4
    assert L.hbs_gemm_streamk(0, P, P, P, 64, 64, 64, 64, 64, 64, ctypes.c_void_p((1 << 20) + 4),
                              need, None) == -3
    big = 0x7fffffff - 4096 + 1
    assert L.hbs_gemm_streamk(0, P, P, P, 8, 8, big, big, 8, 8, W, need, None) == -1
