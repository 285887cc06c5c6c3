"""Split FP32 GEMMs on the 16-bit matrix cores against the extended-precision
oracle.

No reference counterpart: pinned by the oracle and by exactness on inputs the
split represents exactly.  Bars (normwise, vs the oracle):
  BF16, 6 terms: 1e-6   (three BF16 pieces carry 24 bits; the dropped terms are
                         below 2^-24 of each product)
  BF16, 3 terms: 3e-5   (two pieces carry ~17 bits)
  FP16, 3 terms: 1e-6   (scaled hi + 2^-11 lo carry 22 bits)
"""
import numpy as np
import pytest

from conftest import to_dev

pytestmark = pytest.mark.gpu

TOL = {6: 1e-6, 3: 3e-5, "fp16": 1e-6}


def run(orc, blas, dev, mode, M, N, K, nterms, dist=0, alpha=1.5, beta=-0.5):
    import torch

    A = orc.matrix(M, K, dist, 31, dtype=np.float32, ld=(M if mode[0] == "T" else K) + 3, trans=mode[0] == "T")
    B = orc.matrix(K, N, dist, 32, dtype=np.float32, ld=(K if mode[1] == "T" else N) + 1, trans=mode[1] == "T")
    C0 = orc.matrix(M, N, 0, 33, dtype=np.float32, ld=N + 2)
    ref = orc.gemm_xp(mode[0], mode[1], A, B, C0, alpha, beta)
    dA, dB, dC = to_dev(A, dev), to_dev(B, dev), to_dev(C0, dev)
    if nterms == "fp16":
        blas.sgemm_split_fp16(mode[0], mode[1], dA, dB, dC, alpha, beta)
    else:
        blas.sgemm_split_bf16(mode[0], mode[1], dA, dB, dC, alpha, beta, nterms=nterms)
    torch.cuda.synchronize()
    return np.ascontiguousarray(dC.cpu().numpy()), ref


@pytest.mark.parametrize("nterms", [6, 3, "fp16"])
@pytest.mark.parametrize("mode", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("shape", [(1, 1, 1), (33, 17, 70), (128, 128, 64), (200, 136, 300), (257, 129, 1000)])
def test_split_bf16_modes(orc, blas, dev, nterms, mode, shape):
    got, ref = run(orc, blas, dev, mode, *shape, nterms)
    err, _ = orc.normwise(got, ref)
    assert err <= TOL[nterms], err


@pytest.mark.parametrize("scheme", [6, "fp16"])
@pytest.mark.parametrize("dist", [0, 2])
def test_split_accuracy_large_k(orc, blas, dev, dist, scheme):
    """K = 8192 with uniform and log-uniform (2^[-8,8]) magnitudes."""
    got, ref = run(orc, blas, dev, "NN", 64, 96, 8192, scheme, dist=dist, alpha=1.0, beta=0.0)
    err, _ = orc.normwise(got, ref)
    assert err <= 1e-6, err


def test_split_fp16_wide_range_rows(orc, blas, dev):
    """Rows and columns whose magnitudes differ by 2^40 from each other: the
    per-row power-of-two scaling keeps every row in the FP16 range."""
    import torch

    M, N, K = 40, 50, 300
    A = orc.matrix(M, K, 0, 41, dtype=np.float32)
    B = orc.matrix(K, N, 0, 42, dtype=np.float32)
    A = (A * np.float32(2.0) ** np.linspace(-20, 20, M, dtype=np.float32)[:, None]).astype(np.float32)
    B = (B * np.float32(2.0) ** np.linspace(20, -20, N, dtype=np.float32)[None, :]).astype(np.float32)
    C0 = np.zeros((M, N), dtype=np.float32)
    ref = orc.gemm_xp("N", "N", A, B, C0, 1.0, 0.0)
    dC = to_dev(C0, dev)
    blas.sgemm_split_fp16("N", "N", to_dev(A, dev), to_dev(B, dev), dC, 1.0, 0.0)
    torch.cuda.synchronize()
    got = dC.cpu().numpy().astype(np.float64)
    rel = np.abs(got - ref) / np.abs(ref).max(axis=1, keepdims=True)
    assert rel.max() <= 1e-6, rel.max()


def test_split_bf16_exact_on_bf16_inputs(orc, blas, dev):
    """Integer patterns are single BF16 pieces (< 256) and every sum stays
    below 2^24: the result must be exact."""
    import torch

    M, N, K = 130, 140, 1500
    A = orc.matrix(M, K, orc.PATTERN11, 0, dtype=np.float32)
    B = orc.matrix(K, N, orc.PATTERN7, 0, dtype=np.float32)
    P = A.astype(np.int64) @ B.astype(np.int64)
    for nterms in (3, 6, "fp16"):
        C = torch.zeros((M, N), dtype=torch.float32, device=dev)
        if nterms == "fp16":
            blas.sgemm_split_fp16("N", "N", to_dev(A, dev), to_dev(B, dev), C, 1.0, 0.0)
        else:
            blas.sgemm_split_bf16("N", "N", to_dev(A, dev), to_dev(B, dev), C, 1.0, 0.0, nterms=nterms)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(C.cpu().numpy(), P)
