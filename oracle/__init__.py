"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, and only as the checker.  The product path
(hb_hammerbench_amd) never imports this package.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# HBS_ORACLE_LIB selects another build of the same source (the sanitizer
# build, oracle/_asan/liboracle.so, in tests/test_sanitizers.py)
_LIB_PATH = os.environ.get("HBS_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_int = ctypes.c_int


def build():
    """Compile oracle/liboracle.so with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_host_mm": (None, [_f32p, _f32p, _f32p, _int]),
            "orc_host_mm_rows": (None, [_f32p, _f32p, _f32p, _int, _int, _int]),
            "orc_host_mm_omp": (None, [_f32p, _f32p, _f32p, _int, _int, _int, _int]),
            "orc_fmaf_chain_mm": (None, [_f32p, _f32p, _f32p, _int]),
            "orc_pattern_init": (None, [_f32p, _f32p, _i64, _i64]),
            "orc_pattern_value": (_i64, [_int, _int, _int, _int]),
            "orc_pattern_slice": (None, [_f32p, _int, _int]),
            "orc_sse": (ctypes.c_float, [_f32p, _f32p, _i64, ctypes.POINTER(_i64)]),
            "orc_gemm_f64": (None, [_int, _int, _int, _f32p, _int, _f32p, _int, _f64p, _f64p, _int]),
            "orc_errors": (None, [_int, _int, _f32p, _int, _f64p, _f64p, _f64p]),
            "orc_rng64": (_u64, [_u64, _u64]),
            "orc_gen_uniform": (None, [_u64, _u64, _i64, _f32p]),
            "orc_gen_loguniform": (None, [_u64, _u64, _i64, _f32p]),
            "orc_gen_normal_outlier": (None, [_u64, _u64, _i64, _f32p]),
            "orc_bf16_rne": (ctypes.c_uint16, [ctypes.c_float]),
            "orc_bf16_to_float": (ctypes.c_float, [ctypes.c_uint16]),
            "orc_split_bf16": (None, [_f32p, _i64, _u16p, _u16p]),
            "orc_fp8_e4m3_round": (ctypes.c_float, [ctypes.c_float]),
            "orc_fp8_e4m3_encode": (ctypes.c_uint8, [ctypes.c_float]),
            "orc_fp8_e4m3_decode": (ctypes.c_float, [ctypes.c_uint8]),
            "orc_fp8_tile_exponent": (_int, [ctypes.c_float]),
            "orc_fp8_quantize": (None, [_int, _int, _f32p, _int, _int, _f32p, _f32p, _i32p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# --- apps/sgemm/main.cpp:15-26 ------------------------------------------------
def host_mm(A, B):
    A = _c(A, np.float32); B = _c(B, np.float32)
    N = A.shape[0]
    C = np.empty((N, N), np.float32)
    lib().orc_host_mm(C, A, B, N)
    return C


def host_mm_rows(A, B, row_begin, row_end, threads=0):
    """Rows [row_begin, row_end) of host_mm; threads>0 runs the OpenMP variant."""
    A = _c(A, np.float32); B = _c(B, np.float32)
    N = A.shape[0]
    C = np.zeros((N, N), np.float32)
    if threads:
        lib().orc_host_mm_omp(C, A, B, N, row_begin, row_end, threads)
    else:
        lib().orc_host_mm_rows(C, A, B, N, row_begin, row_end)
    return C


# --- apps/sgemm/kernel.cpp:186-275 -------------------------------------------------
def fmaf_chain_mm(A, B):
    A = _c(A, np.float32); B = _c(B, np.float32)
    N = A.shape[0]
    C = np.empty((N, N), np.float32)
    lib().orc_fmaf_chain_mm(C, A, B, N)
    return C


# --- apps/sgemm/main.cpp:46-52 -------------------------------------------------
def pattern_init(n_total, first=0):
    a = np.empty(n_total, np.float32); b = np.empty(n_total, np.float32)
    lib().orc_pattern_init(a, b, first, n_total)
    return a, b


def pattern_value(N, it, y, x):
    return int(lib().orc_pattern_value(N, it, y, x))


def pattern_slice(N, it):
    C = np.empty((N, N), np.float32)
    lib().orc_pattern_slice(C, N, it)
    return C


# --- apps/sgemm/main.cpp:120-134 ------------------------------------------------
def sse(expected, actual):
    e = _c(expected, np.float32).ravel(); a = _c(actual, np.float32).ravel()
    nd = _i64(0)
    s = lib().orc_sse(e, a, e.size, ctypes.byref(nd))
    return float(s), int(nd.value)


# --- FP64 oracle ------------------------------------------------------------------
def gemm_f64(A, B, threads=0):
    """R = A.B and S = |A|.|B| in double (products of floats are exact in double)."""
    A = _c(A, np.float32); B = _c(B, np.float32)
    M, K = A.shape; K2, N = B.shape
    assert K == K2
    R = np.empty((M, N), np.float64)
    S = np.empty((M, N), np.float64)
    lib().orc_gemm_f64(M, N, K, A, K, B, N, R, S, threads)
    return R, S


def errors(C, R, S):
    """(e_cw, e_nw, e_F, max|C-R|, max|R|, n_nonfinite) of C against the FP64 oracle."""
    C = _c(C, np.float32); R = _c(R, np.float64); S = _c(S, np.float64)
    M, N = C.shape
    out = np.zeros(6, np.float64)
    lib().orc_errors(M, N, C, N, R, S, out)
    return dict(e_cw=out[0], e_nw=out[1], e_F=out[2], max_abs_err=out[3],
                max_abs_ref=out[4], nonfinite=int(out[5]))


# --- generators -----------------------------------------------------------------
def rng64(seed, i):
    return int(lib().orc_rng64(seed, i))


def gen(kind, seed, n, first=0):
    out = np.empty(n, np.float32)
    fn = {"uniform": lib().orc_gen_uniform, "loguniform": lib().orc_gen_loguniform,
          "normal_outlier": lib().orc_gen_normal_outlier}[kind]
    fn(seed, first, n, out)
    return out


# --- BF16 / FP8 ------------------------------------------------------------------
def split_bf16(x):
    x = _c(x, np.float32)
    flat = x.ravel()
    hi = np.empty(flat.size, np.uint16); lo = np.empty(flat.size, np.uint16)
    lib().orc_split_bf16(flat, flat.size, hi, lo)
    return hi.reshape(x.shape), lo.reshape(x.shape)


def bf16_bits_to_f32(h):
    return (np.asarray(h, np.uint32) << 16).view(np.float32)


def fp8_round(x):
    return float(lib().orc_fp8_e4m3_round(float(x)))


def fp8_encode(x):
    return int(lib().orc_fp8_e4m3_encode(float(x)))


def fp8_decode(b):
    return float(lib().orc_fp8_e4m3_decode(int(b)))


def fp8_tile_exponent(amax):
    return int(lib().orc_fp8_tile_exponent(float(amax)))


def fp8_quantize(X, tile=128):
    X = _c(X, np.float32)
    rows, cols = X.shape
    Xq = np.empty_like(X); Xr = np.empty_like(X)
    tr = (rows + tile - 1) // tile; tc = (cols + tile - 1) // tile
    e = np.empty(tr * tc, np.int32)
    lib().orc_fp8_quantize(rows, cols, X, cols, tile, Xq, Xr, e)
    return Xq, Xr, e.reshape(tr, tc)
