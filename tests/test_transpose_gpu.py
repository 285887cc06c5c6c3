"""Out-of-place transpose B = A^T (transpose_16x16_dbuf, transpose_float_32x32_dbuf,
transpose_A2B, transpose_A2B_float32, blasdsp_transpose_d/s) against numpy's
transpose of the same host data: bit-exact (a transpose moves values, it does
no arithmetic).  Shapes include the reference harness's cases
(transpose/bin/run.sh: 8210 x 8190, 1512 x 1536, 312 x 228), odd and
one-row/one-column shapes, padded leading dimensions and unaligned bases."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def host(M, N, dtype, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((M, N)).astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(1, 1), (1, 77), (77, 1), (63, 65), (64, 64), (200, 136), (312, 228),
                                   (1512, 1536), (8210, 8190)])
def test_transpose_packed_reference_entry_points(blas, dev, dtype, shape):
    import torch

    M, N = shape
    A = host(M, N, dtype, 1)
    dA = torch.from_numpy(A).to(dev)
    dB = torch.full((N, M), float("nan"), dtype=dA.dtype, device=dev)
    L = blas.lib()
    s = torch.cuda.current_stream().cuda_stream
    fn = L.transpose_16x16_dbuf if dtype == np.float64 else L.transpose_float_32x32_dbuf
    blas.check("transpose", fn(M, N, dA.data_ptr(), dB.data_ptr(), s))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dB.cpu().numpy(), A.T)
    # the transpose_A2B forms
    dB2 = torch.full((N, M), float("nan"), dtype=dA.dtype, device=dev)
    if dtype == np.float64:
        blas.check("transpose_A2B", L.transpose_A2B(N, M, M, N, dA.data_ptr(), dB2.data_ptr(), s))
    else:
        blas.check("transpose_A2B_float32", L.transpose_A2B_float32(M, N, dA.data_ptr(), dB2.data_ptr(), s))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dB2.cpu().numpy(), A.T)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("pad", [(0, 0), (3, 5), (8, 0), (1, 1)])
def test_transpose_leading_dimensions(blas, dev, dtype, pad):
    """Padded A and B; the padding of B is never written."""
    import torch

    M, N = 130, 261
    A = host(M, N + pad[0], dtype, 2)
    dA = torch.from_numpy(A).to(dev)
    dB = torch.full((N, M + pad[1]), 7.0, dtype=dA.dtype, device=dev)
    blas.transpose(dA[:, :N], dB[:, :M])
    torch.cuda.synchronize()
    got = dB.cpu().numpy()
    np.testing.assert_array_equal(got[:, :M], A[:, :N].T)
    assert np.all(got[:, M:] == 7.0)


def test_transpose_unaligned_base(blas, dev):
    """A base one element past a 16-byte boundary takes the element-wise path."""
    import torch

    M, N = 100, 90
    A = host(M, N + 1, np.float32, 3).reshape(-1)
    dflat = torch.from_numpy(A).to(dev)
    dA = dflat[1:1 + M * N].view(M, N)
    dB = torch.empty((N, M), dtype=torch.float32, device=dev)
    blas.transpose(dA, dB)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dB.cpu().numpy(), A[1:1 + M * N].reshape(M, N).T)


def test_transpose_round_trip_large(blas, dev):
    """A^T^T = A at 16384 x 12288 FP32 (768 MiB each way)."""
    import torch

    M, N = 16384, 12288
    dA = torch.empty((M, N), dtype=torch.float32, device=dev)
    blas.fill(dA, M, N, blas.UNIFORM, 5)
    dB = torch.empty((N, M), dtype=torch.float32, device=dev)
    dC = torch.empty((M, N), dtype=torch.float32, device=dev)
    blas.transpose(dA, dB)
    blas.transpose(dB, dC)
    torch.cuda.synchronize()
    assert torch.equal(dA, dC)
    assert torch.equal(dB[:5, :7], dA[:7, :5].T)


def test_transpose_null_and_empty(blas, dev):
    import torch

    L = blas.lib()
    s = torch.cuda.current_stream().cuda_stream
    x = torch.zeros(4, device=dev)
    assert L.transpose_null(2, 2, x.data_ptr(), x.data_ptr(), s) == 0
    assert L.transpose_float_null(2, 2, x.data_ptr(), x.data_ptr(), s) == 0
    assert L.transpose_float_32x32_dbuf(0, 5, None, None, s) == 0
    torch.cuda.synchronize()
