"""Five-point stencil on the GPU: bit-exact against the oracle's restatement of
the reference's host checker (stencil_2d_V1.2/host_code/matrix_stencil.c:33-47)
for packed and haloed outputs, padded and unaligned grids, the reference
harness's shapes (stencil_2d_V1.2/bin/run.sh), the reference-named entry
points and multi-sweep Jacobi iteration."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def grid(rng, M, N, ld=None):
    """Haloed (M+2) x (N+2) grid in a buffer of row length ld (padding NaN)."""
    ld = ld or N + 2
    buf = np.full((M + 2, ld), np.nan)
    buf[:, :N + 2] = rng.uniform(-1.0, 1.0, (M + 2, N + 2))
    return buf


def dev_view(torch, buf, dev, cols, offset=0):
    """Device copy of buf with its row stride; offset shifts the base by that
    many elements (unaligned when odd)."""
    flat = torch.empty(buf.size + offset, dtype=torch.float64, device=dev)
    flat[offset:] = torch.from_numpy(np.ascontiguousarray(buf).ravel()).to(dev)
    return flat[offset:].view(buf.shape)[:, :cols]


@pytest.mark.parametrize("shape", [(1, 1), (1, 7), (7, 1), (3, 2), (55, 111), (33, 513), (64, 1024),
                                   (100, 1023), (257, 2049)])
@pytest.mark.parametrize("halo_out", [False, True])
def test_stencil_bit_exact(orc, blas, dev, shape, halo_out):
    import torch

    rng = np.random.default_rng(shape[0] * 7919 + shape[1])
    M, N = shape
    A = grid(rng, M, N)
    ref = orc.stencil(A, halo_out=halo_out)
    dA = torch.from_numpy(A).to(dev)
    if halo_out:
        dB = torch.full((M + 2, N + 2), float("nan"), dtype=torch.float64, device=dev)
    else:
        dB = torch.full((M, N), float("nan"), dtype=torch.float64, device=dev)
    blas.stencil(dA, dB, halo_out=halo_out)
    torch.cuda.synchronize()
    got = dB.cpu().numpy()
    if halo_out:
        np.testing.assert_array_equal(got[1:-1, 1:-1], ref[1:-1, 1:-1])
        # the halo of B is left as it was
        assert np.isnan(got[0]).all() and np.isnan(got[-1]).all()
        assert np.isnan(got[:, 0]).all() and np.isnan(got[:, -1]).all()
    else:
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("lda_pad,ldb_pad,offset", [(4, 6, 0), (5, 3, 0), (0, 0, 1), (3, 1, 1)])
def test_stencil_padded_and_unaligned(orc, blas, dev, lda_pad, ldb_pad, offset):
    """Leading dimensions past the row (odd ones take the element-wise loads)
    and a base one element off 16-byte alignment; NaN padding is never read."""
    import torch

    rng = np.random.default_rng(11 + lda_pad)
    M, N = 45, 301
    A = grid(rng, M, N, ld=N + 2 + lda_pad)
    ref = orc.stencil(np.ascontiguousarray(A[:, :N + 2]))
    dA = dev_view(torch, A, dev, N + 2, offset)
    Bbuf = np.full((M, N + ldb_pad), np.nan)
    dB = dev_view(torch, Bbuf, dev, N, offset)
    blas.stencil(dA, dB)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dB.cpu().numpy(), ref)


def test_stencil_reference_entry_points(orc, blas, dev):
    """stencil_vector (haloed B, ld N + 2), stencil_scalar (packed B, ld N) and
    stencil_null, with the reference harness's first shape (run.sh: 55 x 111)."""
    import torch

    L = blas.lib()
    rng = np.random.default_rng(3)
    M, N = 55, 111
    A = grid(rng, M, N)
    dA = torch.from_numpy(A).to(dev)
    dH = torch.zeros((M + 2, N + 2), dtype=torch.float64, device=dev)
    dP = torch.zeros((M, N), dtype=torch.float64, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    assert L.stencil_vector(M, N, dA.data_ptr(), dH.data_ptr(), s) == blas.SUCCESS
    assert L.stencil_scalar(M, N, dA.data_ptr(), dP.data_ptr(), s) == blas.SUCCESS
    assert L.stencil_null(M, N, dA.data_ptr(), dP.data_ptr(), s) == blas.SUCCESS
    torch.cuda.synchronize()
    ref = orc.stencil(A)
    np.testing.assert_array_equal(dP.cpu().numpy(), ref)
    np.testing.assert_array_equal(dH.cpu().numpy()[1:-1, 1:-1], ref)


@pytest.mark.parametrize("iters", [1, 2, 5])
def test_stencil_iterate_jacobi(orc, blas, dev, iters):
    """Ping-pong sweeps between two haloed grids (fixed boundary) equal the
    oracle applied iters times, and the result is where the header says."""
    import torch

    rng = np.random.default_rng(20 + iters)
    M, N = 130, 517
    A = grid(rng, M, N)
    want = A.copy()
    for _ in range(iters):
        want = orc.stencil(want, halo_out=True)
    X = torch.from_numpy(A).to(dev)
    Y = X.clone()
    out = blas.stencil_iterate(X, Y, iters)
    torch.cuda.synchronize()
    assert (out is Y) == (iters % 2 == 1)
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_stencil_full_size(orc, blas, dev):
    """The reference harness's largest grid (run.sh: 8192 x 8192), bit-exact
    over the whole interior, and invariant under a 2-row / 3-column shift of
    the window (each output depends only on its five inputs)."""
    import torch

    M = N = 8192
    dA = torch.empty((M + 2, N + 2), dtype=torch.float64, device=dev)
    blas.fill(dA, M + 2, N + 2, blas.UNIFORM, 77)
    dB = torch.empty((M, N), dtype=torch.float64, device=dev)
    blas.stencil(dA, dB)
    sub = torch.empty((M - 2, N - 3), dtype=torch.float64, device=dev)
    blas.stencil(dA[2:, 3:], sub)
    torch.cuda.synchronize()
    A = dA.cpu().numpy()
    got = dB.cpu().numpy()
    np.testing.assert_array_equal(got, orc.stencil(A))
    np.testing.assert_array_equal(sub.cpu().numpy(), got[2:, 3:])
