"""blasdsp_gemm_noalloc, the allocation-free GEMM entry point, against the
entry point it is a variant of (blasdsp_dgemm_ex / blasdsp_sgemm_ex with the
same algo and splitk): the same kernels and K slices, so the same bits, with
the split-K partial sums in the library's static device array instead of a
hipMalloc'd workspace.  Each case runs two calls in a row on one stream, the
second smaller than the first (the array keeps the first call's contents),
and checks both against the existing entry point bit for bit and against the
extended-precision oracle (FP64 1e-13, FP32 (sqrt(K) + 4) 2^-23)."""
import ctypes
import math

import numpy as np
import pytest

from conftest import to_dev

pytestmark = pytest.mark.gpu


def tol(prec, K):
    return 1e-13 if prec == "d" else (math.sqrt(K) + 4) * 2.0 ** -23


def operands(orc, prec, mode, M, N, K, seed):
    dt = np.float64 if prec == "d" else np.float32
    A = orc.matrix(M, K, 0, seed, dtype=dt, ld=(M if mode[0] == "T" else K) + 3, trans=mode[0] == "T")
    B = orc.matrix(K, N, 0, seed + 1, dtype=dt, ld=(K if mode[1] == "T" else N) + 2, trans=mode[1] == "T")
    C = orc.matrix(M, N, 0, seed + 2, dtype=dt, ld=N + 1)
    return A, B, C


def call_noalloc(blas, prec, mode, dA, dB, dC, M, N, K, alpha, beta, algo, splitk, s):
    d = blas.GemmDesc(blas.FLOAT64 if prec == "d" else blas.FLOAT32, mode[0].encode(), mode[1].encode(), M, N, K,
                      alpha, beta, dA.data_ptr(), dA.stride(0), dB.data_ptr(), dB.stride(0), dC.data_ptr(),
                      dC.stride(0))
    return blas.lib().blasdsp_gemm_noalloc(ctypes.byref(d), algo, splitk, s)


def one(orc, blas, dev, prec, mode, M, N, K, algo, splitk, seed):
    import torch

    A, B, C0 = operands(orc, prec, mode, M, N, K, seed)
    alpha, beta = 1.25, -0.5
    dA, dB = to_dev(A, dev), to_dev(B, dev)
    c_ex, c_na = to_dev(C0, dev), to_dev(C0, dev)
    s = torch.cuda.current_stream().cuda_stream
    blas.gemm(mode[0], mode[1], dA, dB, c_ex, alpha, beta, algo=algo, splitk=splitk)
    st = call_noalloc(blas, prec, mode, dA, dB, c_na, M, N, K, alpha, beta, algo, splitk, s)
    assert st == 0, st
    torch.cuda.synchronize()
    ex, na = c_ex.cpu().numpy(), c_na.cpu().numpy()
    np.testing.assert_array_equal(na, ex)
    ref = orc.gemm_xp(mode[0], mode[1], A, B, C0, alpha, beta)
    err, _ = orc.normwise(np.ascontiguousarray(na), ref)
    assert err <= tol(prec, K), err


@pytest.mark.parametrize("prec,mode,first,second,algo,splitk", [
    # tile kernel, explicit K slices: 4 x 1000 x 1000 FP64 partials (32 MB)
    ("d", "NT", (1000, 1000, 4096), (300, 260, 2000), 1, 4),
    ("s", "TN", (1500, 1200, 6000), (129, 77, 1000), 1, 3),
    # skinny kernel, split K (shortA and thinB)
    ("s", "NN", (16, 3000, 8192), (8, 1000, 4096), 2, 8),
    ("d", "TT", (4000, 24, 8192), (3000, 16, 2048), 2, 5),
    # the library's own choice of kernel and slices
    ("d", "NN", (256, 256, 16384), (64, 96, 4096), 0, 0),
    ("s", "TT", (24, 20480, 8192), (16, 3024, 2048), 0, 0),
])
def test_noalloc_matches_allocating_entry(orc, blas, dev, prec, mode, first, second, algo, splitk):
    one(orc, blas, dev, prec, mode, *first, algo, splitk, 11)
    one(orc, blas, dev, prec, mode, *second, algo, splitk, 21)


def test_noalloc_oversized_explicit_split_is_refused(blas, dev):
    """32 slices of 1024 x 1024 FP64 partials need 256 MiB, past the 64 MiB
    array: refused before any launch (C untouched)."""
    import torch

    n = 1024
    A = torch.zeros((n, n), dtype=torch.float64, device=dev)
    C = torch.full((n, n), 7.0, dtype=torch.float64, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    assert call_noalloc(blas, "d", "NN", A, A, C, n, n, n, 1.0, 0.0, blas.ALGO_TILE, 32, s) == blas.ERR_NOMEM
    torch.cuda.synchronize()
    assert bool((C == 7.0).all())
