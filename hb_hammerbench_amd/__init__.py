"""hb_hammerbench_amd -- MI355X (gfx950) implementation of HammerBench's apps/sgemm path.

The product is the C-ABI library ``libhbsgemm.so`` (declared in include/hbsgemm.h)
and the C++ launcher under ``apps/sgemm``.  This module is a thin ctypes binding
used by the tests and bench.py; torch supplies device memory and streams only.

Reference call shape (apps/sgemm/main.cpp:30-149, kernel.cpp:279-333):
``kernel(mat1, mat2, result, pod_id)`` computes ``result_i = mat1_i . mat2_i`` for
NITER row-major N x N slices stored back to back.  :func:`sgemm` is that call.
Errors surface as :class:`HbsError` carrying the library status code, the
counterpart of ``BSG_CUDA_CALL`` returning early on a non-success code.
"""
import ctypes
import os

__all__ = [
    "HbsError", "lib", "library_path", "sgemm", "sgemm_strided", "sgemm_bf16x3",
    "fill_pattern", "fill_random", "verify_f64",
    "GEN_UNIFORM", "GEN_LOGUNIFORM", "GEN_NORMAL_OUTLIER", "PREC_FP32", "PREC_BF16X3",
    "GroupedPlan", "Graph", "gemm_fp8res", "fp8res_workspace_size", "gemm_bf16x3",
    "gemm_fp8res_noalloc", "FP8RES_NOALLOC_SCRATCH_BYTES", "gemm_streamk", "streamk_workspace_size",
    "describe_plan",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libhbsgemm.so"

GEN_UNIFORM, GEN_LOGUNIFORM, GEN_NORMAL_OUTLIER = 0, 1, 2
PREC_FP32, PREC_BF16X3 = 0, 1
TILE_128x128, TILE_64x64, TILE_128x64, TILE_64x128, TILE_256x256 = 0, 1, 2, 3, 4
TILE_128x128_4x1, TILE_128x128_BK32, TILE_128x128_8W, TILE_256x128, TILE_128x256 = 5, 6, 7, 8, 9

_lib = None


class HbsError(RuntimeError):
    """A non-success status code from libhbsgemm."""

    def __init__(self, code, where):
        self.code = code
        msg = lib().hbs_strerror(code).decode()
        super().__init__(f"{where}: status {code}: {msg}")


class ErrStats(ctypes.Structure):
    _fields_ = [("e_cw", ctypes.c_double), ("e_nw", ctypes.c_double),
                ("e_F", ctypes.c_double), ("max_abs_err", ctypes.c_double),
                ("max_abs_ref", ctypes.c_double), ("sum_sq_err", ctypes.c_double),
                ("sum_sq_ref", ctypes.c_double), ("nonfinite", ctypes.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Problem(ctypes.Structure):
    _fields_ = [("A", ctypes.c_void_p), ("B", ctypes.c_void_p), ("C", ctypes.c_void_p),
                ("M", ctypes.c_int), ("N", ctypes.c_int), ("K", ctypes.c_int),
                ("lda", ctypes.c_int), ("ldb", ctypes.c_int), ("ldc", ctypes.c_int)]


class GroupedInfo(ctypes.Structure):
    _fields_ = [("problems", ctypes.c_int), ("workgroups", ctypes.c_int),
                ("tiles", ctypes.c_int64), ("work_items", ctypes.c_int64),
                ("split_tiles", ctypes.c_int64), ("est_imbalance", ctypes.c_double)]


def library_path():
    return os.path.join(_HERE, _LIB_NAME)


_vp, _i, _i64, _u64, _sz = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                            ctypes.c_uint64, ctypes.c_size_t)

# name -> (restype, argtypes); every symbol include/hbsgemm.h declares.
SIGNATURES = {
    "hbs_strerror": (ctypes.c_char_p, [_i]),
    "hbs_version": (_i, []),
    "hbs_sgemm": (_i, [_vp, _vp, _vp, _i, _i, _vp]),
    "hbs_sgemm_strided": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i64, _i64, _i64, _i, _vp]),
    "hbs_sgemm_bf16x3": (_i, [_vp, _vp, _vp, _i, _i, _vp]),
    "hbs_grouped_plan_create": (_i, [ctypes.POINTER(Problem), _i, _i, ctypes.c_char_p,
                                     ctypes.POINTER(_vp)]),
    "hbs_grouped_plan_create_forced": (_i, [ctypes.POINTER(Problem), _i, _i, _i, _i,
                                            ctypes.POINTER(_vp)]),
    "hbs_grouped_plan_info": (_i, [_vp, ctypes.POINTER(GroupedInfo)]),
    "hbs_grouped_plan_describe": (_i, [ctypes.POINTER(Problem), _i, _i, ctypes.c_char_p, _i, _i,
                                       _i, ctypes.POINTER(GroupedInfo)]),
    "hbs_grouped_run": (_i, [_vp, _vp]),
    "hbs_grouped_plan_destroy": (_i, [_vp]),
    "hbs_grouped_graph_create": (_i, [_vp, ctypes.POINTER(_vp)]),
    "hbs_graph_launch": (_i, [_vp, _vp]),
    "hbs_graph_destroy": (_i, [_vp]),
    "hbs_gemm_bf16x3": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _vp]),
    "hbs_gemm_fp8res_workspace_size": (_sz, [_i, _i, _i]),
    "hbs_gemm_fp8res": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _vp, _sz, _vp]),
    "hbs_gemm_fp8res_noalloc": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _vp]),
    "hbs_gemm_streamk_workspace_size": (_sz, [_i, _i, _i, _i]),
    "hbs_gemm_streamk": (_i, [_i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _vp, _sz, _vp]),
    "hbs_fill_pattern": (_i, [_vp, _vp, _i64, _i64, _vp]),
    "hbs_fill_random": (_i, [_vp, _i64, _i, _u64, _u64, _vp]),
    "hbs_verify_f64": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _i,
                            ctypes.POINTER(ErrStats), _vp]),
}


def lib():
    """Load the in-tree libhbsgemm.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: run __graft_entry__.build() "
                              "(make -C hb_hammerbench_amd/csrc)")
        L = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


_planner = None


def planner():
    """The library that runs the grouped planner's dry run: libhbsgemm.so, or
    the host-only build named by HBS_PLANNER_LIB (the sanitizer build of the
    planner, hb_hammerbench_amd/_asan/libhbsplan_asan.so)."""
    global _planner
    if _planner is None:
        path = os.environ.get("HBS_PLANNER_LIB")
        if path:
            P = ctypes.CDLL(path)
            res, args = SIGNATURES["hbs_grouped_plan_describe"]
            P.hbs_grouped_plan_describe.restype = res
            P.hbs_grouped_plan_describe.argtypes = args
            _planner = P
        else:
            _planner = lib()
    return _planner


def _check(code, where):
    if code != 0:
        raise HbsError(code, where)


def _stream(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return getattr(stream, "cuda_stream", stream)


def _ptr(t):
    if t is None:
        return None
    return t.data_ptr() if hasattr(t, "data_ptr") else t


def sgemm(mat1, mat2, result, N, niter=1, stream=None):
    """result_i = mat1_i . mat2_i for i < niter N x N row-major slices (kernel.cpp:279-333)."""
    _check(lib().hbs_sgemm(_ptr(mat1), _ptr(mat2), _ptr(result), N, niter, _stream(stream)),
           "hbs_sgemm")


def sgemm_strided(A, B, C, M, N, K, lda, ldb, ldc, strideA=0, strideB=0, strideC=0, batch=1,
                  stream=None):
    _check(lib().hbs_sgemm_strided(_ptr(A), _ptr(B), _ptr(C), M, N, K, lda, ldb, ldc,
                                   strideA, strideB, strideC, batch, _stream(stream)),
           "hbs_sgemm_strided")


def sgemm_bf16x3(mat1, mat2, result, N, niter=1, stream=None):
    """BF16x3 split-precision SGEMM with the hbs_sgemm call shape (no workspace)."""
    _check(lib().hbs_sgemm_bf16x3(_ptr(mat1), _ptr(mat2), _ptr(result), N, niter,
                                  _stream(stream)),
           "hbs_sgemm_bf16x3")


def gemm_bf16x3(A, B, C, M, N, K, lda=None, ldb=None, ldc=None, stream=None):
    """Rectangular BF16x3 split-precision GEMM (no workspace)."""
    _check(lib().hbs_gemm_bf16x3(_ptr(A), _ptr(B), _ptr(C), M, N, K,
                                 K if lda is None else lda, N if ldb is None else ldb,
                                 N if ldc is None else ldc, _stream(stream)),
           "hbs_gemm_bf16x3")


def fp8res_workspace_size(M, N, K):
    return int(lib().hbs_gemm_fp8res_workspace_size(M, N, K))


def gemm_fp8res(A, B, C, M, N, K, lda=None, ldb=None, ldc=None, workspace=None, stream=None):
    """FP8-e4m3 (per-128x128-tile power-of-two scales) + BF16 residual GEMM."""
    ws_bytes = fp8res_workspace_size(M, N, K)
    if workspace is None:
        import torch
        workspace = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device="cuda")
    nbytes = workspace.numel() * workspace.element_size()
    _check(lib().hbs_gemm_fp8res(_ptr(A), _ptr(B), _ptr(C), M, N, K,
                                 K if lda is None else lda, N if ldb is None else ldb,
                                 N if ldc is None else ldc, _ptr(workspace), nbytes,
                                 _stream(stream)),
           "hbs_gemm_fp8res")
    return workspace


# Bytes of the fixed scratch of hbs_gemm_fp8res_noalloc (HBS_FP8RES_NOALLOC_SCRATCH_BYTES).
FP8RES_NOALLOC_SCRATCH_BYTES = 64 << 20


def gemm_fp8res_noalloc(A, B, C, M, N, K, lda=None, ldb=None, ldc=None, stream=None):
    """gemm_fp8res with the workspace in the library's fixed scratch: no device
    allocation (problems whose workspace exceeds FP8RES_NOALLOC_SCRATCH_BYTES
    raise HbsError with HBS_ERR_SCRATCH).  Calls on one GPU share that scratch
    and must not overlap."""
    _check(lib().hbs_gemm_fp8res_noalloc(_ptr(A), _ptr(B), _ptr(C), M, N, K,
                                         K if lda is None else lda, N if ldb is None else ldb,
                                         N if ldc is None else ldc, _stream(stream)),
           "hbs_gemm_fp8res_noalloc")


def streamk_workspace_size(M, N, K, precision=PREC_FP32):
    return int(lib().hbs_gemm_streamk_workspace_size(precision, M, N, K))


def gemm_streamk(A, B, C, M, N, K, precision=PREC_FP32, lda=None, ldb=None, ldc=None,
                 workspace=None, stream=None):
    """Stream-K FP32 or BF16x3 GEMM.  Without `workspace` a zero-filled one is
    allocated; a workspace passed in must have been zero-filled once (every
    call leaves its counters zero).  Returns the workspace for reuse."""
    ws_bytes = streamk_workspace_size(M, N, K, precision)
    if workspace is None:
        import torch
        workspace = torch.zeros(max(ws_bytes, 1), dtype=torch.uint8, device="cuda")
    nbytes = workspace.numel() * workspace.element_size()
    _check(lib().hbs_gemm_streamk(precision, _ptr(A), _ptr(B), _ptr(C), M, N, K,
                                  K if lda is None else lda, N if ldb is None else ldb,
                                  N if ldc is None else ldc, _ptr(workspace), nbytes,
                                  _stream(stream)),
           "hbs_gemm_streamk")
    return workspace


def fill_pattern(mat1, mat2, n, first=0, stream=None):
    """mat1[i] = i % 7, mat2[i] = i % 3 (apps/sgemm/main.cpp:46-52)."""
    _check(lib().hbs_fill_pattern(_ptr(mat1), _ptr(mat2), n, first, _stream(stream)),
           "hbs_fill_pattern")


def fill_random(dst, n, kind, seed, first=0, stream=None):
    _check(lib().hbs_fill_random(_ptr(dst), n, kind, seed, first, _stream(stream)),
           "hbs_fill_random")


def verify_f64(A, B, C, M, N, K, lda=None, ldb=None, ldc=None, row_begin=0, row_end=None,
               stream=None):
    """FP64 error statistics of C against A.B computed on the device (synchronous)."""
    out = ErrStats()
    _check(lib().hbs_verify_f64(_ptr(A), _ptr(B), _ptr(C), M, N, K,
                                K if lda is None else lda, N if ldb is None else ldb,
                                N if ldc is None else ldc, row_begin,
                                M if row_end is None else row_end, ctypes.byref(out),
                                _stream(stream)),
           "hbs_verify_f64")
    return out.as_dict()


def describe_plan(shapes, precision=PREC_FP32, tuning_file=None, tile=None, split_k=1, cus=256):
    """Counts of the grouped plan for problems of these (M, N, K) shapes on a GPU
    of `cus` compute units, without a device (hbs_grouped_plan_describe).
    Operands are taken as contiguous (lda = K, ldb = N, ldc = N); the pointers
    are placeholders that are never dereferenced."""
    arr = (Problem * len(shapes))()
    fake = 1 << 20  # a 16-byte-aligned placeholder address
    for i, (M, N, K) in enumerate(shapes):
        arr[i] = Problem(fake, fake, fake, M, N, K, K, N, N)
    out = GroupedInfo()
    path = tuning_file.encode() if tuning_file else None
    _check(planner().hbs_grouped_plan_describe(arr, len(shapes), precision, path,
                                           -1 if tile is None else tile, split_k, cus,
                                           ctypes.byref(out)), "hbs_grouped_plan_describe")
    return {name: getattr(out, name) for name, _ in GroupedInfo._fields_}


class GroupedPlan:
    """A grouped persistent GEMM plan over (A, B, C, M, N, K[, lda, ldb, ldc]) problems.

    Tensors are referenced, not copied: keep them alive while the plan is used.
    """

    def __init__(self, problems, precision=PREC_FP32, tuning_file=None, tile=None, split_k=1):
        arr = (Problem * len(problems))()
        self._keep = []
        for i, pr in enumerate(problems):
            A, B, C, M, N, K = pr[:6]
            lda, ldb, ldc = (pr[6:9] if len(pr) >= 9 else (K, N, N))
            self._keep += [A, B, C]
            arr[i] = Problem(_ptr(A), _ptr(B), _ptr(C), M, N, K, lda, ldb, ldc)
        h = ctypes.c_void_p()
        if tile is None:
            path = tuning_file.encode() if tuning_file else None
            rc = lib().hbs_grouped_plan_create(arr, len(problems), precision, path,
                                               ctypes.byref(h))
        else:
            rc = lib().hbs_grouped_plan_create_forced(arr, len(problems), precision, tile,
                                                      split_k, ctypes.byref(h))
        _check(rc, "hbs_grouped_plan_create")
        self.handle = h
        self.precision = precision

    def info(self):
        out = GroupedInfo()
        _check(lib().hbs_grouped_plan_info(self.handle, ctypes.byref(out)), "hbs_grouped_plan_info")
        return {name: getattr(out, name) for name, _ in GroupedInfo._fields_}

    def run(self, stream=None):
        _check(lib().hbs_grouped_run(self.handle, _stream(stream)), "hbs_grouped_run")

    def graph(self):
        return Graph(self)

    def close(self):
        if self.handle:
            _check(lib().hbs_grouped_plan_destroy(self.handle), "hbs_grouped_plan_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """An executable hipGraph holding one run of a GroupedPlan."""

    def __init__(self, plan):
        h = ctypes.c_void_p()
        _check(lib().hbs_grouped_graph_create(plan.handle, ctypes.byref(h)),
               "hbs_grouped_graph_create")
        self.handle = h
        self.plan = plan

    def launch(self, stream=None):
        _check(lib().hbs_graph_launch(self.handle, _stream(stream)), "hbs_graph_launch")

    def close(self):
        if self.handle:
            _check(lib().hbs_graph_destroy(self.handle), "hbs_graph_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
