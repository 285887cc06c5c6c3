"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker (or, for cpu_baseline, as
the timed CPU restatement of the reference host GEMM).  The product library
never imports it.

All arrays are numpy arrays in row-major storage; a matrix with leading
dimension ld is a 2-D array whose row stride is ld elements (a column slice of
a wider array, or a contiguous array with ld == number of columns).
"""
from __future__ import annotations

import ctypes
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

UNIFORM, NORMAL, LOGUNIF, PATTERN11, PATTERN7, ONES = range(6)

_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_dp = ctypes.c_void_p
_lib = None


def build(force: bool = False) -> None:
    args = ["make", "-C", str(HERE)]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True, capture_output=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.oracle_mix64.restype = _u64
        L.oracle_mix64.argtypes = [_u64]
        L.oracle_value.restype = ctypes.c_double
        L.oracle_value.argtypes = [ctypes.c_int, _u64, _u64]
        L.oracle_fill_f64.argtypes = [_dp, _i64, _i64, _i64, ctypes.c_int, ctypes.c_int, _u64, ctypes.c_double]
        L.oracle_fill_f32.argtypes = [_dp, _i64, _i64, _i64, ctypes.c_int, ctypes.c_int, _u64, ctypes.c_float]
        L.oracle_dgemm_v11_host.argtypes = [_i64, _i64, _i64, _dp, _i64, _dp, _i64, _dp, _i64, _i64]
        L.oracle_sgemm_plain_host.argtypes = [_i64, _i64, _i64, _dp, _i64, _dp, _i64, _dp, _i64, _i64]
        L.oracle_sgemm_v132_host.argtypes = [_i64, _i64, _i64, ctypes.c_float, ctypes.c_float,
                                             _dp, _i64, _dp, _i64, _dp, _i64, _i64]
        L.oracle_thinB_sgemm_v12_host.argtypes = [_i64, _i64, _i64, ctypes.c_float, ctypes.c_float,
                                                  _dp, _i64, _dp, _i64, _dp, _i64, _i64, _i64]
        L.oracle_dgemm_blas_host.argtypes = [_i64, _i64, _i64, ctypes.c_double, _dp, _i64, _dp, _i64,
                                             ctypes.c_double, _dp, _i64, _i64]
        L.oracle_sgemm_blas_host.argtypes = [_i64, _i64, _i64, ctypes.c_float, _dp, _i64, _dp, _i64,
                                             ctypes.c_float, _dp, _i64, _i64]
        L.oracle_gemm_xp_f64.argtypes = [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, ctypes.c_double,
                                         _dp, _i64, _dp, _i64, ctypes.c_double, _dp, _i64]
        L.oracle_gemm_xp_f32.argtypes = [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, ctypes.c_double,
                                         _dp, _i64, _dp, _i64, ctypes.c_double, _dp, _i64, _dp, _i64]
        L.oracle_normwise_f64.restype = ctypes.c_double
        L.oracle_normwise_f64.argtypes = [_i64, _i64, _dp, _i64, _dp, _i64, ctypes.POINTER(ctypes.c_double)]
        L.oracle_normwise_f32.restype = ctypes.c_double
        L.oracle_normwise_f32.argtypes = [_i64, _i64, _dp, _i64, _dp, _i64, ctypes.POINTER(ctypes.c_double)]
        for f in (L.oracle_project_f64, L.oracle_project_f32):
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_char, _i64, _i64, _i64, _dp, _i64, _dp, _dp]
        L.oracle_fp8_e4m3.restype = ctypes.c_uint
        L.oracle_fp8_e4m3.argtypes = [ctypes.c_float]
        L.oracle_fp8_decode.restype = ctypes.c_float
        L.oracle_fp8_decode.argtypes = [ctypes.c_uint]
        L.oracle_fp8_block_exp.restype = ctypes.c_int
        L.oracle_fp8_block_exp.argtypes = [ctypes.c_float]
        L.oracle_sgemm_fp8.argtypes = [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, ctypes.c_double,
                                       _dp, _i64, _dp, _i64, ctypes.c_double, _dp, _i64, _dp, _i64]
        L.oracle_stencil.argtypes = [_i64, _i64, _dp, _i64, _dp, _i64, ctypes.c_int]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _ld(a: np.ndarray) -> int:
    assert a.ndim == 2
    if a.shape[0] <= 1 or a.shape[1] == 0:
        return max(a.shape[1], 1)
    assert a.strides[1] == a.itemsize or a.shape[1] == 1, "rows must be contiguous"
    return a.strides[0] // a.itemsize


def set_threads(n: int) -> None:
    lib().oracle_set_threads(int(n))


def threads() -> int:
    return int(lib().oracle_get_threads())


def value(dist: int, seed: int, idx: int) -> float:
    return float(lib().oracle_value(dist, seed, idx))


def matrix(rows: int, cols: int, dist: int, seed: int, dtype=np.float64, ld: int | None = None,
           trans: bool = False, pad: float = np.nan) -> np.ndarray:
    """Logical rows x cols matrix from the seeded generator.

    Returns the storage view: rows x cols (trans=False) or cols x rows
    (trans=True, i.e. the transposed copy), with row stride ld; padding
    columns hold `pad` (NaN by default so that reading them shows up)."""
    srows, scols = (cols, rows) if trans else (rows, cols)
    ld = scols if ld is None else ld
    assert ld >= max(1, scols)
    buf = np.empty((max(srows, 1), max(ld, 1)), dtype=dtype)
    fn = lib().oracle_fill_f64 if dtype == np.float64 else lib().oracle_fill_f32
    fn(_ptr(buf), rows, cols, ld, int(trans), dist, seed, pad)
    return buf[:srows, :scols]


def dgemm_v11_host(A, B, C, K_G: int = 512) -> None:
    M, K = A.shape
    N = B.shape[1]
    lib().oracle_dgemm_v11_host(M, N, K, _ptr(A), _ld(A), _ptr(B), _ld(B), _ptr(C), _ld(C), K_G)


def sgemm_plain_host(A, B, C, K_G: int = 512) -> None:
    M, K = A.shape
    N = B.shape[1]
    lib().oracle_sgemm_plain_host(M, N, K, _ptr(A), _ld(A), _ptr(B), _ld(B), _ptr(C), _ld(C), K_G)


def sgemm_v132_host(A, B, C, alpha: float, beta: float, K_G: int = 512) -> None:
    M, K = A.shape
    N = B.shape[1]
    lib().oracle_sgemm_v132_host(M, N, K, alpha, beta, _ptr(A), _ld(A), _ptr(B), _ld(B), _ptr(C), _ld(C), K_G)


def thinB_sgemm_v12_host(A, B, C, alpha: float, beta: float, K_G: int = 3072, K_A: int = 512) -> None:
    M, K = A.shape
    N = B.shape[1]
    lib().oracle_thinB_sgemm_v12_host(M, N, K, alpha, beta, _ptr(A), _ld(A), _ptr(B), _ld(B), _ptr(C),
                                      _ld(C), K_G, K_A)


def gemm_blas_host(A, B, C, alpha: float, beta: float, K_G: int = 512) -> None:
    """The CPU baseline: reference accumulation tree + BLAS epilogue (NN)."""
    M, K = A.shape
    N = B.shape[1]
    fn = lib().oracle_dgemm_blas_host if A.dtype == np.float64 else lib().oracle_sgemm_blas_host
    fn(M, N, K, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B), beta, _ptr(C), _ld(C), K_G)


def gemm_xp(transA: str, transB: str, A, B, C, alpha: float, beta: float) -> np.ndarray:
    """Extended-precision C = alpha op(A) op(B) + beta C; returns a new
    float64 M x N array (C is not modified)."""
    M, N = C.shape
    K = A.shape[0] if transA == "T" else A.shape[1]
    out = np.empty((M, N), dtype=np.float64)
    if A.dtype == np.float64:
        out[...] = C
        lib().oracle_gemm_xp_f64(transA.encode(), transB.encode(), M, N, K, alpha, _ptr(A), _ld(A),
                                 _ptr(B), _ld(B), beta, _ptr(out), N)
    else:
        lib().oracle_gemm_xp_f32(transA.encode(), transB.encode(), M, N, K, alpha, _ptr(A), _ld(A),
                                 _ptr(B), _ld(B), beta, _ptr(C), _ld(C), _ptr(out), N)
    return out


def project(trans: str, A, X) -> np.ndarray:
    """op(A) X with extended-precision sums, rounded once to float64: A float32
    or float64 with contiguous rows (op(A) = A^T when trans == "T"), X a
    float64 K x S array with 1 <= S <= 16."""
    M = A.shape[1] if trans == "T" else A.shape[0]
    K = A.shape[0] if trans == "T" else A.shape[1]
    X = np.ascontiguousarray(X, dtype=np.float64)
    assert X.shape[0] == K and 1 <= X.shape[1] <= 16, (X.shape, K)
    Y = np.empty((M, X.shape[1]), dtype=np.float64)
    fn = lib().oracle_project_f64 if A.dtype == np.float64 else lib().oracle_project_f32
    assert A.dtype in (np.float64, np.float32)
    rc = fn(trans.encode(), M, K, X.shape[1], _ptr(A), _ld(A), _ptr(X), _ptr(Y))
    assert rc == 0, rc
    return Y


def gemm_projection(transA: str, A, transB: str, B, C0, alpha: float, beta: float, X,
                    side: str = "right") -> np.ndarray:
    """The random projection of R = alpha op(A) op(B) + beta C0, without forming R:
    side "right": R X = alpha op(A) (op(B) X) + beta C0 X (X is N x S, result M x S);
    side "left": R^T X = alpha op(B)^T (op(A)^T X) + beta C0^T X (X is M x S, result N x S)."""
    flip = {"N": "T", "T": "N"}
    if side == "right":
        RX = alpha * project(transA, A, project(transB, B, X))
        c = ("N", C0)
    else:
        RX = alpha * project(flip[transB], B, project(flip[transA], A, X))
        c = ("T", C0)
    if beta != 0.0:
        RX += beta * project(c[0], c[1], X)
    return RX


def projected_errors(G, RX, X, side: str = "right") -> tuple[float, float, int]:
    """Error estimates for a whole GEMM result G from its random projection
    (Freivalds' method): RX is gemm_projection's for the same X and side, X has
    independent N(0, 1) entries.  A line is a row of G (side "right") or a
    column (side "left").  With E = G - R, E[||E X||_F^2] = S ||E||_F^2 and
    E[||(E X)_i||^2] = S ||E_i||^2 for line i, so returns
    (||E X||_F / ||R X||_F, the normwise relative error's estimate;
    max_i ||(E X)_i|| / (sqrt(S) ||G_i||), the worst line's relative error
    estimate; that line's index).  An error spread over many lines is
    estimated to a few percent; a rank-one error (one k term lost everywhere)
    only to the factor chi_S / sqrt(S), within [0.4, 1.8] but for odds below
    1e-3 at S = 8."""
    S = X.shape[1]
    right = side == "right"
    E = project("N" if right else "T", G, X) - RX
    den = np.linalg.norm(RX)
    nw = float(np.linalg.norm(E) / den) if den > 0 else (0.0 if not E.any() else np.inf)
    g = np.sqrt(np.einsum("ij,ij->i" if right else "ij,ij->j", G, G, dtype=np.float64))
    with np.errstate(invalid="ignore", divide="ignore"):  # a non-finite G gives non-finite estimates
        line = np.linalg.norm(E, axis=1) / (np.sqrt(S) * np.maximum(g, np.finfo(np.float64).tiny))
    i = int(np.argmax(line))
    return nw, float(line[i]), i


def gemm_fp8(transA: str, transB: str, A, B, C, alpha: float, beta: float) -> np.ndarray:
    """The scaled-FP8 GEMM of float32 operands (quantized as the library
    quantizes them, accumulated in extended precision); returns a new float64
    M x N array."""
    M, N = C.shape
    K = A.shape[0] if transA == "T" else A.shape[1]
    out = np.empty((M, N), dtype=np.float64)
    lib().oracle_sgemm_fp8(transA.encode(), transB.encode(), M, N, K, alpha, _ptr(A), _ld(A),
                           _ptr(B), _ld(B), beta, _ptr(C), _ld(C), _ptr(out), N)
    return out


def fp8_e4m3(x: float) -> int:
    return int(lib().oracle_fp8_e4m3(x))


def fp8_decode(q: int) -> float:
    return float(lib().oracle_fp8_decode(q))


def fp8_block_exp(amax: float) -> int:
    return int(lib().oracle_fp8_block_exp(amax))


def normwise(X, R) -> tuple[float, float]:
    """(||X - R||_F / ||R||_F, max |X - R|) with R float64."""
    M, N = R.shape
    mx = ctypes.c_double(0.0)
    fn = lib().oracle_normwise_f64 if X.dtype == np.float64 else lib().oracle_normwise_f32
    err = fn(M, N, _ptr(X), _ld(X), _ptr(R), _ld(R), ctypes.byref(mx))
    return float(err), float(mx.value)


def stencil(A: np.ndarray, halo_out: bool = False) -> np.ndarray:
    """Five-point stencil of a haloed float64 grid A ((M+2) x (N+2) view, rows
    contiguous): the packed M x N result, or with halo_out the (M+2) x (N+2)
    grid whose interior holds it (halo copied from A)."""
    assert A.dtype == np.float64
    M, N = A.shape[0] - 2, A.shape[1] - 2
    if halo_out:
        B = np.array(A, dtype=np.float64, copy=True)
    else:
        B = np.empty((M, N), dtype=np.float64)
    lib().oracle_stencil(M, N, _ptr(A), _ld(A), _ptr(B), _ld(B), 1 if halo_out else 0)
    return B
