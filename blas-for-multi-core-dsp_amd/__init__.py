"""Python binding of the blasdsp C ABI (include/blasdsp.h) over ctypes.

The product is the C library lib/libblasdsp.so; this module only loads it,
declares the argument types of every exported function and offers
tensor-level helpers for tests and the benchmark.  Device memory and streams
come from PyTorch (plumbing only): a torch tensor's data pointer and the
current HIP stream are handed to the C functions unchanged.

There is no fallback: if the library is missing or fails to load, every call
raises.  Build it with `make -C blas-for-multi-core-dsp_amd` (or
__graft_entry__.build()).
"""
from __future__ import annotations

import ctypes
import pathlib
import re

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "lib" / "libblasdsp.so"
HEADER = HERE.parent / "include" / "blasdsp.h"

SUCCESS = 0
ERR_INVALID = -22
ERR_NOMEM = -12
ERR_DEVICE = -5
ERR_UNSUPPORTED = -95

ALGO_AUTO, ALGO_TILE, ALGO_SKINNY = 0, 1, 2

UNIFORM, NORMAL, LOGUNIF, PATTERN11, PATTERN7, ONES = range(6)

_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_d = ctypes.c_double
_f = ctypes.c_float
_lib = None


FLOAT32, FLOAT64 = 0, 1


class GemmDesc(ctypes.Structure):
    """blasdsp_gemm_desc (include/blasdsp.h)."""
    _fields_ = [("dtype", ctypes.c_int), ("transA", ctypes.c_char), ("transB", ctypes.c_char),
                ("M", ctypes.c_int64), ("N", ctypes.c_int64), ("K", ctypes.c_int64),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
                ("A", ctypes.c_void_p), ("lda", ctypes.c_int64),
                ("B", ctypes.c_void_p), ("ldb", ctypes.c_int64),
                ("C", ctypes.c_void_p), ("ldc", ctypes.c_int64)]


class Request(ctypes.Structure):
    """blasdsp_request."""
    _fields_ = [("dtype", ctypes.c_int), ("transA", ctypes.c_char), ("transB", ctypes.c_char),
                ("M", ctypes.c_int64), ("N", ctypes.c_int64), ("K", ctypes.c_int64),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("arrival_s", ctypes.c_double)]


class RequestResult(ctypes.Structure):
    """blasdsp_request_result."""
    _fields_ = [("start_s", ctypes.c_double), ("done_s", ctypes.c_double), ("latency_s", ctypes.c_double),
                ("checksum", ctypes.c_double), ("batch", ctypes.c_int), ("status", ctypes.c_int)]


class ServeConfig(ctypes.Structure):
    """blasdsp_serve_config."""
    _fields_ = [("memory_budget", ctypes.c_int64), ("max_batch", ctypes.c_int), ("max_inflight", ctypes.c_int),
                ("grouped", ctypes.c_int), ("time_scale", ctypes.c_double)]


class ServeStats(ctypes.Structure):
    """blasdsp_serve_stats."""
    _fields_ = [("completed", ctypes.c_int64), ("rejected", ctypes.c_int64), ("wall_s", ctypes.c_double),
                ("gemms_per_s", ctypes.c_double), ("tflops", ctypes.c_double), ("p50_s", ctypes.c_double),
                ("p99_s", ctypes.c_double), ("mean_s", ctypes.c_double), ("max_s", ctypes.c_double),
                ("launches", ctypes.c_int64), ("mean_batch", ctypes.c_double),
                ("peak_resident_bytes", ctypes.c_int64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class TuneChoice(ctypes.Structure):
    """blasdsp_tune_choice."""
    _fields_ = [("algo", ctypes.c_int), ("tile", ctypes.c_int), ("splitk", ctypes.c_int),
                ("time_us", ctypes.c_double)]


class BlasdspError(RuntimeError):
    def __init__(self, fn: str, status: int):
        super().__init__(f"{fn} returned {status}")
        self.status = status


def header_functions() -> list[str]:
    """Names of every function declared in include/blasdsp.h."""
    text = HEADER.read_text()
    names = re.findall(r"^BLASDSP_API\s+[\w\s\*]*?\b(\w+)\s*\(", text, flags=re.M)
    names += re.findall(r"^BLASDSP_DECLARE_MODE\((\w+),", text, flags=re.M)
    return names


def _declare(L: ctypes.CDLL) -> None:
    mode_args = [_i64, _i64, _i64, _i64, _i64, _i64, _vp, _vp, _vp]
    for name in header_functions():
        fn = getattr(L, name)
        fn.restype = ctypes.c_int
    L.blasdsp_version.restype = ctypes.c_char_p
    L.blasdsp_status_string.restype = ctypes.c_char_p
    L.blasdsp_status_string.argtypes = [ctypes.c_int]
    gen = [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64]
    L.blasdsp_dgemm.argtypes = gen + [_d, _vp, _i64, _vp, _i64, _d, _vp, _i64, _vp]
    L.blasdsp_sgemm.argtypes = gen + [_f, _vp, _i64, _vp, _i64, _f, _vp, _i64, _vp]
    L.blasdsp_dgemm_ex.argtypes = gen + [_d, _vp, _i64, _vp, _i64, _d, _vp, _i64, ctypes.c_int,
                                         ctypes.c_int, _vp]
    L.blasdsp_sgemm_ex.argtypes = gen + [_f, _vp, _i64, _vp, _i64, _f, _vp, _i64, ctypes.c_int,
                                         ctypes.c_int, _vp]
    packed = [_i64, _i64, _i64, _vp, _vp, _vp]
    for name in header_functions():
        if re.match(r"^(dgemm|blasdsp_irr_\w+_dgemm)_(NN|NT|TN|TT|null)$", name):
            getattr(L, name).argtypes = mode_args + [_d, _d, _vp]
        elif re.match(r"^blasdsp_(sgemm|irr_\w+_sgemm)_(NN|NT|TN|TT)$", name):
            getattr(L, name).argtypes = mode_args + [_f, _f, _vp]
        elif re.match(r"^(sgemm_\w\w_qxx|sgemm_null|irr_thinB_sgemm_(NN|NT|TN|TT|null))$", name):
            getattr(L, name).argtypes = packed + [_f, _f, _vp]
        elif re.match(r"^irr_(shortA|thinB)_[ds]gemm_\w+$", name):
            getattr(L, name).argtypes = packed + [_vp]
    L.blasdsp_group_exec.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_uint64), _vp]
    L.blasdsp_gemm_noalloc.argtypes = [ctypes.POINTER(GemmDesc), ctypes.c_int, ctypes.c_int, _vp]
    L.blasdsp_fill_f64.argtypes = [_vp, _i64, _i64, _i64, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _d, _vp]
    L.blasdsp_fill_f32.argtypes = [_vp, _i64, _i64, _i64, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _f, _vp]
    L.blasdsp_sgemm_split_bf16.argtypes = gen + [_f, _vp, _i64, _vp, _i64, _f, _vp, _i64, ctypes.c_int, _vp]
    L.blasdsp_sgemm_split_fp16.argtypes = gen + [_f, _vp, _i64, _vp, _i64, _f, _vp, _i64, _vp]
    L.blasdsp_sgemm_fp8.argtypes = gen + [_f, _vp, _i64, _vp, _i64, _f, _vp, _i64, _vp]
    L.blasdsp_gemm_grouped.argtypes = [ctypes.POINTER(GemmDesc), ctypes.c_int, _vp]
    L.blasdsp_group_plan_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    L.blasdsp_group_plan_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(GemmDesc), ctypes.c_int, _vp]
    L.blasdsp_group_plan_destroy.argtypes = [ctypes.c_void_p]
    L.blasdsp_serve.argtypes = [ctypes.POINTER(Request), ctypes.c_int, ctypes.POINTER(ServeConfig),
                                ctypes.POINTER(RequestResult), ctypes.POINTER(ServeStats),
                                ctypes.POINTER(ctypes.c_void_p)]
    tune_args = [ctypes.c_int, ctypes.c_char, ctypes.c_char, _i64, _i64, _i64]
    L.blasdsp_autotune.argtypes = tune_args + [_vp, ctypes.POINTER(TuneChoice)]
    L.blasdsp_tune_query.argtypes = tune_args + [ctypes.POINTER(TuneChoice)]
    L.blasdsp_tune_save.argtypes = [ctypes.c_char_p]
    L.blasdsp_tune_load.argtypes = [ctypes.c_char_p]
    L.blasdsp_tune_clear.argtypes = []
    L.blasdsp_probe_read.argtypes = [_vp, _i64, _vp, ctypes.c_int, _vp]
    L.blasdsp_probe_mfma.argtypes = [ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]
    L.blasdsp_probe_lds.argtypes = [ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]
    L.blasdsp_stream_release.argtypes = [_vp]
    for name in ("transpose_16x16_dbuf", "transpose_float_32x32_dbuf", "transpose_null", "transpose_float_null",
                 "transpose_A2B_float32"):
        getattr(L, name).argtypes = [_i64, _i64, _vp, _vp, _vp]
    L.transpose_A2B.argtypes = [_i64, _i64, _i64, _i64, _vp, _vp, _vp]
    L.blasdsp_transpose_d.argtypes = [_i64, _i64, _vp, _i64, _vp, _i64, _vp]
    L.blasdsp_transpose_s.argtypes = [_i64, _i64, _vp, _i64, _vp, _i64, _vp]
    for name in ("stencil_vector", "stencil_scalar", "stencil_null"):
        getattr(L, name).argtypes = [_i64, _i64, _vp, _vp, _vp]
    L.blasdsp_stencil.argtypes = [_i64, _i64, _vp, _i64, _vp, _i64, ctypes.c_int, _vp]
    L.blasdsp_stencil_iterate.argtypes = [_i64, _i64, _vp, _vp, _i64, ctypes.c_int, _vp]
    for extra in _EXTRA_DECLS:
        extra(L)


_EXTRA_DECLS: list = []


def _bind_hip_runtime() -> None:
    """Load PyTorch's HIP runtime before the library when torch is installed.

    PyTorch ships its own libamdhip64 (soname libamdhip64.so.7).  Loaded
    first, it satisfies the library's libamdhip64.so.7 dependency, so one HIP
    runtime serves the process.  Loaded after the library has pulled in the
    system copy, torch's loader asks for "libamdhip64.so", finds no loaded
    object of that name and maps a second HIP (and HSA) runtime beside the
    first."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib() -> ctypes.CDLL:
    """Load lib/libblasdsp.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is not built; run make -C {HERE}")
        _bind_hip_runtime()
        L = ctypes.CDLL(str(LIB_PATH))
        _declare(L)
        _lib = L
    return _lib


def check(fn: str, status: int) -> None:
    if status != SUCCESS:
        raise BlasdspError(fn, status)


# ---- tensor helpers ---------------------------------------------------------

def _stream(stream):
    import torch

    if stream is None:
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


def _ld(t) -> int:
    if t.dim() != 2 or (t.stride(1) != 1 and t.shape[1] > 1):
        raise ValueError("matrices must be 2-D with contiguous rows")
    return int(t.stride(0)) if t.shape[0] > 1 else max(int(t.shape[1]), 1)


def gemm(transA: str, transB: str, A, B, C, alpha: float = 1.0, beta: float = 0.0,
         algo: int = ALGO_AUTO, splitk: int = 0, stream=None, raise_on_error: bool = True) -> int:
    """C = alpha op(A) op(B) + beta C on torch device tensors (in place on C).

    A, B, C are the STORED matrices (A is K x M for transA == 'T').  Leading
    dimensions are the tensors' row strides, so column slices of wider
    buffers express padded leading dimensions."""
    import torch

    M, N = C.shape
    K = A.shape[0] if transA.upper() == "T" else A.shape[1]
    if C.dtype == torch.float64:
        fn, name = lib().blasdsp_dgemm_ex, "blasdsp_dgemm_ex"
    elif C.dtype == torch.float32:
        fn, name = lib().blasdsp_sgemm_ex, "blasdsp_sgemm_ex"
    else:
        raise TypeError(f"unsupported dtype {C.dtype}")
    assert A.dtype == B.dtype == C.dtype
    st = fn(transA.encode(), transB.encode(), M, N, K, alpha, A.data_ptr(), _ld(A), B.data_ptr(), _ld(B),
            beta, C.data_ptr(), _ld(C), algo, splitk, _stream(stream))
    if raise_on_error:
        check(name, st)
    return st


def sgemm_split_bf16(transA: str, transB: str, A, B, C, alpha: float = 1.0, beta: float = 0.0,
                     nterms: int = 6, stream=None) -> None:
    """Split-BF16 FP32 GEMM on torch float32 device tensors (in place on C)."""
    M, N = C.shape
    K = A.shape[0] if transA.upper() == "T" else A.shape[1]
    check("blasdsp_sgemm_split_bf16",
          lib().blasdsp_sgemm_split_bf16(transA.encode(), transB.encode(), M, N, K, alpha, A.data_ptr(), _ld(A),
                                         B.data_ptr(), _ld(B), beta, C.data_ptr(), _ld(C), nterms,
                                         _stream(stream)))


def desc(transA: str, transB: str, A, B, C, alpha: float = 1.0, beta: float = 0.0) -> GemmDesc:
    """Descriptor of C = alpha op(A) op(B) + beta C over torch device tensors."""
    import torch

    M, N = C.shape
    K = A.shape[0] if transA.upper() == "T" else A.shape[1]
    dt = FLOAT64 if C.dtype == torch.float64 else FLOAT32
    return GemmDesc(dt, transA.encode(), transB.encode(), M, N, K, alpha, beta, A.data_ptr(), _ld(A),
                    B.data_ptr(), _ld(B), C.data_ptr(), _ld(C))


def desc_array(descs):
    """A ctypes GemmDesc array of `descs`, as blasdsp_gemm_grouped and
    blasdsp_group_plan_run take it."""
    return (GemmDesc * len(descs))(*descs)


def gemm_grouped(descs, stream=None) -> None:
    """All problems of `descs` (GemmDesc list) in one grouped launch."""
    arr = (GemmDesc * len(descs))(*descs)
    check("blasdsp_gemm_grouped", lib().blasdsp_gemm_grouped(arr, len(descs), _stream(stream)))


class GroupPlan:
    """blasdsp_group_plan: a reusable hipGraph of the grouped kernels."""

    def __init__(self, max_problems: int):
        self._p = ctypes.c_void_p()
        check("blasdsp_group_plan_create", lib().blasdsp_group_plan_create(max_problems, ctypes.byref(self._p)))

    def run(self, descs, stream=None) -> None:
        """descs: a list of GemmDesc, or a ctypes GemmDesc array built once
        (desc_array) and reused across runs."""
        arr = descs if isinstance(descs, ctypes.Array) else desc_array(descs)
        check("blasdsp_group_plan_run", lib().blasdsp_group_plan_run(self._p, arr, len(arr), _stream(stream)))

    def close(self) -> None:
        if self._p:
            check("blasdsp_group_plan_destroy", lib().blasdsp_group_plan_destroy(self._p))
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve(requests, config: ServeConfig, outputs=None):
    """Run blasdsp_serve over a list of Request; returns (results, stats).
    outputs: None or a list of host buffers (numpy arrays) or None per request."""
    n = len(requests)
    reqs = (Request * n)(*requests)
    res = (RequestResult * n)()
    stats = ServeStats()
    outp = None
    if outputs is not None:
        outp = (ctypes.c_void_p * n)(*[None if o is None else o.ctypes.data for o in outputs])
    check("blasdsp_serve", lib().blasdsp_serve(reqs, n, ctypes.byref(config), res, ctypes.byref(stats), outp))
    return list(res), stats


def autotune(dtype: int, transA: str, transB: str, M: int, N: int, K: int, stream=None) -> TuneChoice:
    c = TuneChoice()
    check("blasdsp_autotune", lib().blasdsp_autotune(dtype, transA.encode(), transB.encode(), M, N, K,
                                                     _stream(stream), ctypes.byref(c)))
    return c


def tune_query(dtype: int, transA: str, transB: str, M: int, N: int, K: int):
    """The recorded choice for the shape class, or None."""
    c = TuneChoice()
    st = lib().blasdsp_tune_query(dtype, transA.encode(), transB.encode(), M, N, K, ctypes.byref(c))
    return c if st == SUCCESS else None


def sgemm_split_fp16(transA: str, transB: str, A, B, C, alpha: float = 1.0, beta: float = 0.0,
                     stream=None) -> None:
    """Scaled split-FP16 FP32 GEMM on torch float32 device tensors (in place on C)."""
    M, N = C.shape
    K = A.shape[0] if transA.upper() == "T" else A.shape[1]
    check("blasdsp_sgemm_split_fp16",
          lib().blasdsp_sgemm_split_fp16(transA.encode(), transB.encode(), M, N, K, alpha, A.data_ptr(), _ld(A),
                                         B.data_ptr(), _ld(B), beta, C.data_ptr(), _ld(C), _stream(stream)))


def sgemm_fp8(transA: str, transB: str, A, B, C, alpha: float = 1.0, beta: float = 0.0, stream=None) -> None:
    """Scaled-FP8 (E4M3, per-128-block scales) FP32 GEMM on torch float32 device tensors (in place on C)."""
    M, N = C.shape
    K = A.shape[0] if transA.upper() == "T" else A.shape[1]
    check("blasdsp_sgemm_fp8",
          lib().blasdsp_sgemm_fp8(transA.encode(), transB.encode(), M, N, K, alpha, A.data_ptr(), _ld(A),
                                  B.data_ptr(), _ld(B), beta, C.data_ptr(), _ld(C), _stream(stream)))


def transpose(A, B, stream=None) -> None:
    """B = A^T on torch device tensors (float32 or float64; 2-D, unit column stride)."""
    import torch

    M, N = A.shape
    if B.shape != (N, M) or A.dtype != B.dtype:
        raise ValueError("B must be A.shape[::-1] with A's dtype")
    fn = lib().blasdsp_transpose_d if A.dtype == torch.float64 else lib().blasdsp_transpose_s
    check("blasdsp_transpose", fn(M, N, A.data_ptr(), _ld(A), B.data_ptr(), _ld(B), _stream(stream)))


def stencil(A, B, halo_out: bool = False, stream=None) -> None:
    """One five-point stencil sweep of the haloed float64 grid A ((M+2) x
    (N+2) device tensor, unit column stride) into B: (M+2) x (N+2) with
    halo_out (interior written), else M x N."""
    import torch

    M, N = A.shape[0] - 2, A.shape[1] - 2
    want = (M + 2, N + 2) if halo_out else (M, N)
    if A.dtype != torch.float64 or B.dtype != torch.float64 or tuple(B.shape) != want:
        raise ValueError(f"A float64 (M+2) x (N+2), B float64 {want}")
    check("blasdsp_stencil", lib().blasdsp_stencil(M, N, A.data_ptr(), _ld(A), B.data_ptr(), _ld(B),
                                                   int(halo_out), _stream(stream)))


def stencil_iterate(X, Y, iters: int, stream=None):
    """`iters` Jacobi sweeps between the haloed grids X and Y (same shape and
    row stride); returns the tensor holding the last sweep's result."""
    M, N = X.shape[0] - 2, X.shape[1] - 2
    if X.shape != Y.shape or _ld(X) != _ld(Y):
        raise ValueError("X and Y must have the same shape and row stride")
    check("blasdsp_stencil_iterate", lib().blasdsp_stencil_iterate(M, N, X.data_ptr(), Y.data_ptr(), _ld(X),
                                                                   iters, _stream(stream)))
    return Y if iters % 2 else X


def fill(X, rows: int, cols: int, dist: int, seed: int, trans: bool = False, pad: float = 0.0, stream=None):
    """Device generator into the storage tensor X (rows x cols logical)."""
    import torch

    fn = lib().blasdsp_fill_f64 if X.dtype == torch.float64 else lib().blasdsp_fill_f32
    check("blasdsp_fill", fn(X.data_ptr(), rows, cols, int(X.stride(0)), int(trans), dist, seed, pad,
                             _stream(stream)))
    return X


def version() -> str:
    return lib().blasdsp_version().decode()
