"""GPU parity: the HIP kernels, called through the C ABI, against the CPU oracle.

Tolerances (normwise ||C - C_ref||_F / ||C_ref||_F against the
extended-precision oracle):
  FP64: 1e-13.
  FP32: (sqrt(K) + 4) * 2^-23 — sqrt(K) * eps32 for the K-term dot products plus
        a few ulps for the alpha/beta epilogue.
Integer-pattern inputs (the reference demos' %11 / %7 matrices) are exact in
both precisions and are compared bit for bit.

Full-size results are checked whole through the oracle's random projection
(Freivalds' method, S = 8 Gaussian vectors): the normwise estimate against the
bars above, and every row's (or column's) relative error estimate against
1e-12 (FP64) or 4 (sqrt(K) + 4) 2^-23 (FP32) — about 2x for the spread of
the largest of many chi-square estimates, 2x for rows with cancellation.
"""
import math

import numpy as np
import pytest

from conftest import to_dev

pytestmark = pytest.mark.gpu

TOL64 = 1e-13
LINE64 = 1e-12
PROJ_S = 8


def tol32(K):
    return (math.sqrt(max(K, 1)) + 4) * 2.0 ** -23


def line32(K):
    return 4 * tol32(K)


def dtype_np(prec):
    return np.float64 if prec == "d" else np.float32


def stored(orc, rows, cols, trans, dist, seed, prec, pad=3):
    """op(X) is rows x cols; returns the stored (maybe transposed) padded view."""
    scols = rows if trans else cols
    return orc.matrix(rows, cols, dist, seed, dtype=dtype_np(prec), ld=scols + pad, trans=trans)


def run_case(orc, blas, dev, prec, ta, tb, M, N, K, alpha, beta, algo, dist=0, seeds=(1, 2, 3), splitk=0,
             pad=3, cpad=5):
    """Stored A and B with leading dimension (stored columns + pad), C with
    N + cpad: the default pads leave rows unaligned (bounds-checked paths);
    pad = cpad = 4 aligns every row to 16 bytes (the unmasked paths)."""
    import torch

    A = stored(orc, M, K, ta == "T", dist, seeds[0], prec, pad=pad)
    B = stored(orc, K, N, tb == "T", dist, seeds[1], prec, pad=pad)
    C0 = orc.matrix(M, N, dist if dist in (3, 4) else 0, seeds[2], dtype=dtype_np(prec), ld=N + cpad)
    ref = orc.gemm_xp(ta, tb, A, B, C0, alpha, beta)
    dA, dB, dC = to_dev(A, dev), to_dev(B, dev), to_dev(C0, dev)
    blas.gemm(ta, tb, dA, dB, dC, alpha, beta, algo=algo, splitk=splitk)
    torch.cuda.synchronize()
    got = np.ascontiguousarray(dC.cpu().numpy())
    return got, ref


SMALL_SHAPES = [(1, 1, 1), (7, 5, 3), (16, 16, 4), (17, 33, 65), (128, 128, 16), (129, 130, 131),
                (200, 136, 300), (64, 300, 97)]


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("mode", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("shape", SMALL_SHAPES)
def test_tile_kernel_modes(orc, blas, dev, prec, mode, shape):
    M, N, K = shape
    got, ref = run_case(orc, blas, dev, prec, mode[0], mode[1], M, N, K, 1.5, -0.5, blas.ALGO_TILE)
    err, _ = orc.normwise(got, ref)
    assert err <= (TOL64 if prec == "d" else tol32(K)), err


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("mode", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("shape", [(8, 300, 1000), (16, 64, 16), (32, 1000, 777), (40, 129, 64),
                                   (1, 70, 33), (300, 8, 1000), (1000, 32, 777), (129, 40, 64), (70, 1, 33),
                                   (32, 100, 20), (300, 20, 5), (20, 1030, 1000), (1030, 20, 1000),
                                   (48, 700, 1500), (1100, 33, 900)])
@pytest.mark.parametrize("pad", [3, 4])
def test_skinny_kernel_modes(orc, blas, dev, prec, mode, shape, pad):
    """Every mode against the oracle, with rows unaligned (pad 3: bounds-checked
    loads) and 16-byte aligned (pad 4: unmasked buffer loads for whole panels
    and K blocks, 16-byte stores of C).  With K = 1000 the FP32 32-row K-split
    waves of the 20 x 1030 shapes walk 16 and 15 blocks."""
    M, N, K = shape
    got, ref = run_case(orc, blas, dev, prec, mode[0], mode[1], M, N, K, 1.5, -0.5, blas.ALGO_SKINNY,
                        pad=pad, cpad=pad + 2 if pad == 3 else pad)
    err, _ = orc.normwise(got, ref)
    assert err <= (TOL64 if prec == "d" else tol32(K)), err


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("mode", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("shape", [(300, 260, 256), (513, 257, 320), (256, 384, 197)])
def test_tile_alignment_invariance(orc, blas, dev, prec, mode, shape):
    """The tile kernel with operand rows 16-byte aligned (pad 4) and not (pads
    1 and 3): a vector is one 16-byte load at any element alignment, edges and
    the K tail read as zero through the buffer range check, so C is bit for bit
    the same, and within the bar of the oracle.  K = 256 and 320 give interior
    tiles a whole number of K tiles (the unchecked body), K = 197 a K tail."""
    M, N, K = shape
    outs = []
    for pad in (4, 1, 3):
        got, ref = run_case(orc, blas, dev, prec, mode[0], mode[1], M, N, K, 1.5, -0.5, blas.ALGO_TILE,
                            splitk=1, pad=pad, cpad=pad)
        err, _ = orc.normwise(got, ref)
        assert err <= (TOL64 if prec == "d" else tol32(K)), (pad, err)
        outs.append(got)
    np.testing.assert_array_equal(outs[0], outs[1])
    np.testing.assert_array_equal(outs[0], outs[2])


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("mode", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("splitk", [2, 3, 5])
@pytest.mark.parametrize("shape", [(200, 136, 1000), (257, 129, 1111), (64, 64, 2048), (130, 300, 515)])
def test_tile_splitk(orc, blas, dev, prec, mode, splitk, shape):
    """Tile kernel with K cut into slices (the last one ragged when K is not a
    whole number of K tiles), partials added by the reduction kernel."""
    M, N, K = shape
    got, ref = run_case(orc, blas, dev, prec, mode[0], mode[1], M, N, K, 1.5, -0.5, blas.ALGO_TILE,
                        splitk=splitk)
    err, _ = orc.normwise(got, ref)
    assert err <= (TOL64 if prec == "d" else tol32(K)), err


def test_tile_splitk_deterministic_and_auto(orc, blas, dev):
    """Split-K results are bit-reproducible, and the automatic choice (which
    splits K for this shape: 25 tiles of 64 on a 256-CU GPU) matches the
    oracle."""
    import torch

    M, N, K = 320, 320, 4096
    A = orc.matrix(M, K, 0, 81)
    B = orc.matrix(K, N, 0, 82)
    C0 = orc.matrix(M, N, 0, 83)
    ref = orc.gemm_xp("N", "N", A, B, C0, 1.0, 1.0)
    outs = []
    for splitk in (0, 0, 4, 4):
        dC = to_dev(C0, dev)
        blas.gemm("N", "N", to_dev(A, dev), to_dev(B, dev), dC, 1.0, 1.0, splitk=splitk)
        torch.cuda.synchronize()
        outs.append(dC.cpu().numpy())
        err, _ = orc.normwise(outs[-1], ref)
        assert err <= TOL64, (splitk, err)
    np.testing.assert_array_equal(outs[0], outs[1])
    np.testing.assert_array_equal(outs[2], outs[3])


@pytest.mark.parametrize("splitk", [1, 2, 3, 7])
def test_skinny_splitk(orc, blas, dev, splitk):
    for prec in "ds":
        got, ref = run_case(orc, blas, dev, prec, "N", "N", 24, 520, 1111, 1.0, 1.0, blas.ALGO_SKINNY,
                            splitk=splitk)
        err, _ = orc.normwise(got, ref)
        assert err <= (TOL64 if prec == "d" else tol32(1111)), (prec, err)


# Short long sides take many K slices: with splitk 0 the library's picker
# fills one round of workgroups (up to K / 512 slices, here 32 and 40), and
# explicit counts up to 64 cut K into slices shorter than the 512 minimum.
@pytest.mark.parametrize("case", [("d", "NN", 40, 700, 16384, 0), ("s", "TN", 700, 24, 16384, 0),
                                  ("d", "NT", 24, 520, 16384, 64), ("s", "NN", 16, 300, 20480, 0),
                                  ("s", "TT", 300, 48, 20480, 57)])
def test_skinny_many_slices(orc, blas, dev, case):
    prec, mode, M, N, K, splitk = case
    got, ref = run_case(orc, blas, dev, prec, mode[0], mode[1], M, N, K, 1.25, -0.75, blas.ALGO_SKINNY,
                        splitk=splitk, pad=4, cpad=4)
    err, _ = orc.normwise(got, ref)
    assert err <= (TOL64 if prec == "d" else tol32(K)), (case, err)


@pytest.mark.parametrize("prec", ["d", "s"])
@pytest.mark.parametrize("splitk", [1, 3])
@pytest.mark.parametrize("case", [("NN", 32, 1100, 4000), ("TT", 1100, 32, 4000), ("NN", 28, 2100, 1000),
                                  ("TT", 2100, 28, 1000), ("TN", 32, 1100, 4000), ("TN", 1100, 32, 4000),
                                  ("TN", 28, 2100, 1000), ("TN", 2100, 28, 1000), ("TN", 30, 1100, 1000),
                                  ("NN", 16, 1100, 4000), ("TT", 1100, 16, 4000), ("TN", 12, 2100, 1000),
                                  ("TN", 2100, 16, 1000)])
def test_skinny_shared_small_operand(orc, blas, dev, prec, splitk, case):
    """16- and 32-row skinny calls with aligned rows and a k-major streamed operand:
    whole workgroups of full panels share the small operand through LDS in
    chunks of blocks, read K-contiguous (shortA NN, thinB TT) or m-contiguous
    (shortA TN, thinB TN; 30 rows is not a whole number of FP32 vectors, so
    FP32 takes the per-wave path there).  K =
    4000 and 1000 give several chunks with a partial last one, a ragged last
    K block, and with 3 slices a different chunk walk per slice; the partial
    last workgroup (long side 1100 / 2100) takes the per-wave path.  Rows
    16-byte aligned (unmasked loads, vector stores of C).  Against the
    oracle."""
    mode, M, N, K = case
    got, ref = run_case(orc, blas, dev, prec, mode[0], mode[1], M, N, K, 1.5, -0.5, blas.ALGO_SKINNY,
                        splitk=splitk, pad=4, cpad=4)
    err, _ = orc.normwise(got, ref)
    assert err <= (TOL64 if prec == "d" else tol32(K)), err


@pytest.mark.parametrize("algo", ["tile", "skinny"])
@pytest.mark.parametrize("mode", ["NN", "NT", "TN", "TT"])
def test_integer_patterns_exact(orc, blas, dev, algo, mode):
    """The reference demos' integer matrices: every partial sum is an integer
    below 2^24, so both precisions must reproduce the exact product."""
    a = blas.ALGO_TILE if algo == "tile" else blas.ALGO_SKINNY
    M, N, K = (32, 200, 1300) if algo == "skinny" else (150, 130, 1300)
    for prec in "ds":
        A = stored(orc, M, K, mode[0] == "T", orc.PATTERN11, 0, prec)
        B = stored(orc, K, N, mode[1] == "T", orc.PATTERN7, 0, prec)
        P = (orc.matrix(M, K, orc.PATTERN11, 0).astype(np.int64)
             @ orc.matrix(K, N, orc.PATTERN7, 0).astype(np.int64))
        C0 = orc.matrix(M, N, orc.PATTERN7, 5, dtype=dtype_np(prec))
        import torch
        dA, dB, dC = to_dev(A, dev), to_dev(B, dev), to_dev(C0, dev)
        blas.gemm(mode[0], mode[1], dA, dB, dC, 5.0, 3.0, algo=a)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(dC.cpu().numpy(), 5 * P + 3 * C0.astype(np.int64))


def test_reference_entry_points(orc, blas, dev):
    """Per-mode entry points with reference names and argument order; the
    reference's C += op(A)op(B) is alpha = beta = 1."""
    import torch

    L = blas.lib()
    M, N, K = 96, 80, 700
    P = (orc.matrix(M, K, orc.PATTERN11, 0).astype(np.int64)
         @ orc.matrix(K, N, orc.PATTERN7, 0).astype(np.int64))
    s = torch.cuda.current_stream().cuda_stream
    for prec, T in (("d", "dgemm"), ("s", "sgemm")):
        for mode in ("NN", "NT", "TN", "TT"):
            for fam in ("", "irr_shortA_", "irr_thinB_"):
                A = to_dev(stored(orc, M, K, mode[0] == "T", orc.PATTERN11, 0, prec), dev)
                B = to_dev(stored(orc, K, N, mode[1] == "T", orc.PATTERN7, 0, prec), dev)
                C = to_dev(orc.matrix(M, N, orc.PATTERN7, 5, dtype=dtype_np(prec), ld=N + 1), dev)
                C_in = C.cpu().numpy().astype(np.int64)
                fn = getattr(L, f"dgemm_{mode}" if fam == "" and T == "dgemm" else f"blasdsp_{fam}{T}_{mode}")
                st = fn(A.stride(0), B.stride(0), C.stride(0), M, N, K, A.data_ptr(), B.data_ptr(),
                        C.data_ptr(), 1.0, 1.0, s)
                assert st == 0
                torch.cuda.synchronize()
                np.testing.assert_array_equal(C.cpu().numpy(), P + C_in, err_msg=f"{fam}{T}_{mode}")
    assert L.dgemm_null(K, N, N, M, N, K, A.data_ptr(), B.data_ptr(), C.data_ptr(), 1.0, 1.0, s) == 0


def test_beta_zero_ignores_c_and_degenerate_k(orc, blas, dev):
    import torch

    for prec, dt in (("d", torch.float64), ("s", torch.float32)):
        A = to_dev(stored(orc, 40, 50, False, 0, 1, prec), dev)
        B = to_dev(stored(orc, 50, 60, False, 0, 2, prec), dev)
        for algo in (blas.ALGO_TILE, blas.ALGO_SKINNY):
            C = torch.full((40, 60), float("nan"), dtype=dt, device=dev)
            blas.gemm("N", "N", A, B, C, 1.0, 0.0, algo=algo)
            torch.cuda.synchronize()
            assert torch.isfinite(C).all()
        # K = 0 and alpha = 0: C = beta C
        C0 = torch.arange(40 * 60, dtype=dt, device=dev).reshape(40, 60)
        C = C0.clone()
        blas.gemm("N", "N", A[:, :0], B[:0, :], C, 1.0, 2.0)
        torch.cuda.synchronize()
        assert torch.equal(C, 2 * C0)
        C = C0.clone()
        blas.gemm("N", "N", A, B, C, 0.0, -1.0)
        torch.cuda.synchronize()
        assert torch.equal(C, -C0)
        C = torch.full((40, 60), float("nan"), dtype=dt, device=dev)
        blas.gemm("N", "N", A, B, C, 0.0, 0.0)
        torch.cuda.synchronize()
        assert torch.equal(C, torch.zeros_like(C))


def test_fill_matches_host_generator(orc, blas, dev):
    import torch

    for dist in (orc.UNIFORM, orc.PATTERN11, orc.PATTERN7, orc.ONES):
        for trans in (False, True):
            ref = orc.matrix(33, 47, dist, 9, ld=(33 if trans else 47) + 2, trans=trans, pad=0.0)
            X = torch.empty(ref.base.shape, dtype=torch.float64, device=dev)
            blas.fill(X, 33, 47, dist, 9, trans=trans, pad=0.0)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(X.cpu().numpy(), ref.base)
            X32 = torch.empty(ref.base.shape, dtype=torch.float32, device=dev)
            blas.fill(X32, 33, 47, dist, 9, trans=trans, pad=0.0)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(X32.cpu().numpy(), ref.base.astype(np.float32))
    for dist in (orc.NORMAL, orc.LOGUNIF):
        ref = orc.matrix(64, 64, dist, 4)
        X = torch.empty((64, 64), dtype=torch.float64, device=dev)
        blas.fill(X, 64, 64, dist, 4)
        torch.cuda.synchronize()
        np.testing.assert_allclose(X.cpu().numpy(), ref, rtol=1e-13, atol=1e-15)


def test_dgemm_4096_all_modes(orc, blas, dev):
    """DGEMM 4096^3 with lda = ldb = ldc = 4104, alpha = 1.5, beta = -0.5,
    U[-1,1] from seed 1: 32 rows of every mode against the oracle, every row
    of every mode through the oracle's random projection, and the four modes
    against each other over the whole matrix."""
    import torch

    n, ld = 4096, 4104
    alpha, beta = 1.5, -0.5
    outs = {}
    for mode in ("NN", "NT", "TN", "TT"):
        A = torch.empty((n, ld), dtype=torch.float64, device=dev)
        B = torch.empty((n, ld), dtype=torch.float64, device=dev)
        C = torch.empty((n, ld), dtype=torch.float64, device=dev)
        blas.fill(A, n, n, 0, 11, trans=mode[0] == "T")
        blas.fill(B, n, n, 0, 12, trans=mode[1] == "T")
        blas.fill(C, n, n, 0, 13)
        blas.gemm(mode[0], mode[1], A[:, :n], B[:, :n], C[:, :n], alpha, beta)
        torch.cuda.synchronize()
        outs[mode] = C[:, :n].cpu().numpy()
        del A, B, C
    r0, nr = 1000, 32
    A_full = orc.matrix(n, n, 0, 11)
    B_full = orc.matrix(n, n, 0, 12)
    C_full = orc.matrix(n, n, 0, 13)
    ref = orc.gemm_xp("N", "N", np.ascontiguousarray(A_full[r0:r0 + nr]), B_full,
                      np.ascontiguousarray(C_full[r0:r0 + nr]), alpha, beta)
    X = np.random.default_rng(41).standard_normal((n, PROJ_S))
    RX = orc.gemm_projection("N", A_full, "N", B_full, C_full, alpha, beta, X)
    for mode, C in outs.items():
        err, _ = orc.normwise(np.ascontiguousarray(C[r0:r0 + nr]), ref)
        assert err <= TOL64, (mode, err)
        nw, row, i = orc.projected_errors(C, RX, X)
        print(f"dgemm 4096 {mode}: projected normwise {nw:.2e}, worst row {row:.2e} (row {i})")
        assert nw <= TOL64 and row <= LINE64, (mode, nw, row, i)
        merr, _ = orc.normwise(C, outs["NN"])
        assert merr <= 1e-15, (mode, merr)


@pytest.mark.parametrize("prec", ["s", "d"])
@pytest.mark.parametrize("kind", ["shortA", "thinB"])
def test_skinny_full_size(orc, blas, dev, prec, kind):
    """shortA M = 32, K = 8192, N = 65536 (thinB: M and N swapped), N(0,1) from
    seed 2, alpha = beta = 1: 96 columns (rows) against the oracle, the whole
    result through the oracle's random projection, and the automatic split-K
    result against a single-slice run."""
    import torch

    dt = torch.float64 if prec == "d" else torch.float32
    small, K, large = 32, 8192, 65536
    M, N = (small, large) if kind == "shortA" else (large, small)
    A = torch.empty((M, K), dtype=dt, device=dev)
    B = torch.empty((K, N), dtype=dt, device=dev)
    C = torch.empty((M, N), dtype=dt, device=dev)
    blas.fill(A, M, K, 1, 21)
    blas.fill(B, K, N, 1, 22)
    blas.fill(C, M, N, 1, 23)
    j0, nj = 4000, 96
    band = (slice(None), slice(j0, j0 + nj)) if kind == "shortA" else (slice(j0, j0 + nj), slice(None))
    c0 = np.ascontiguousarray(C[band].cpu().numpy())
    c0_full = C.cpu().numpy()
    C1 = C.clone()
    blas.gemm("N", "N", A, B, C, 1.0, 1.0)
    blas.gemm("N", "N", A, B, C1, 1.0, 1.0, algo=blas.ALGO_SKINNY, splitk=1)
    torch.cuda.synchronize()
    tol = TOL64 if prec == "d" else tol32(K)
    if kind == "shortA":
        a = np.ascontiguousarray(A.cpu().numpy())
        b = np.ascontiguousarray(B[:, j0:j0 + nj].cpu().numpy())
    else:
        a = np.ascontiguousarray(A[j0:j0 + nj].cpu().numpy())
        b = np.ascontiguousarray(B.cpu().numpy())
    got = np.ascontiguousarray(C[band].cpu().numpy())
    ref = orc.gemm_xp("N", "N", a, b, c0, 1.0, 1.0)
    err, _ = orc.normwise(got, ref)
    assert err <= tol, err
    got_full = C.cpu().numpy()
    serr, _ = orc.normwise(got_full, C1.cpu().numpy().astype(np.float64))
    assert serr <= tol, serr
    # the whole result: every long line (rows of shortA's C, columns of
    # thinB's) through the projection from the long side
    side = "right" if kind == "shortA" else "left"
    X = np.random.default_rng(42).standard_normal((large, PROJ_S))
    RX = orc.gemm_projection("N", A.cpu().numpy(), "N", B.cpu().numpy(), c0_full, 1.0, 1.0, X, side=side)
    nw, line, i = orc.projected_errors(got_full, RX, X, side=side)
    print(f"{prec} {kind}: projected normwise {nw:.2e}, worst line {line:.2e} ({i})")
    assert nw <= tol and line <= (LINE64 if prec == "d" else line32(K)), (nw, line, i)


@pytest.mark.parametrize("ldb", [20_000_000, 23_000_000])
def test_skinny_wide_leading_dimension(orc, blas, dev, ldb):
    """shortA FP32 (M = 8, K = 32, N = 96) with B a 96-column window of rows
    ldb elements apart, through ALGO_SKINNY.  Below about 22.4 million
    elements the skinny kernel's per-lane byte offsets (up to 96 rows of B)
    stay under 2^31 and it runs the GEMM; above, the call falls back to the
    tile kernel.  Both against the oracle."""
    import torch

    M, N, K = 8, 96, 32
    A = torch.empty((M, K), dtype=torch.float32, device=dev)
    Bbuf = torch.empty((K, ldb), dtype=torch.float32, device=dev)
    B = Bbuf[:, :N]
    C = torch.empty((M, N), dtype=torch.float32, device=dev)
    blas.fill(A, M, K, 0, 31)
    blas.fill(B, K, N, 0, 32)
    blas.fill(C, M, N, 0, 33)
    c0 = np.ascontiguousarray(C.cpu().numpy())
    blas.gemm("N", "N", A, B, C, 1.5, -0.5, algo=blas.ALGO_SKINNY)
    torch.cuda.synchronize()
    a = np.ascontiguousarray(A.cpu().numpy())
    b = np.ascontiguousarray(B.cpu().numpy())
    ref = orc.gemm_xp("N", "N", a, b, c0, 1.5, -0.5)
    err, _ = orc.normwise(np.ascontiguousarray(C.cpu().numpy()), ref)
    del Bbuf
    torch.cuda.empty_cache()
    assert err <= tol32(K), err


def test_stream_release_frees_scratch(orc, blas, dev):
    """A split-K GEMM and a grouped call give a stream scratch memory and a
    cached plan; blasdsp_stream_release frees them (twice is harmless), and
    the stream works again afterwards."""
    import torch

    s = torch.cuda.Stream()
    M, N, K = 512, 512, 8192
    A = torch.rand((M, K), dtype=torch.float64, device=dev)
    B = torch.rand((K, N), dtype=torch.float64, device=dev)
    ref = (A.cpu() @ B.cpu()).numpy()
    C = torch.zeros((M, N), dtype=torch.float64, device=dev)
    Cg = torch.zeros((64, 64), dtype=torch.float64, device=dev)
    with torch.cuda.stream(s):
        blas.gemm("N", "N", A, B, C, 1.0, 0.0, algo=blas.ALGO_TILE, splitk=16, stream=s.cuda_stream)
        blas.gemm_grouped([blas.desc("N", "N", A[:64, :64], B[:64, :64], Cg, 1.0, 0.0)], stream=s.cuda_stream)
    s.synchronize()
    assert np.allclose(C.cpu().numpy(), ref, rtol=1e-12)
    assert np.allclose(Cg.cpu().numpy(), (A[:64, :64].cpu() @ B[:64, :64].cpu()).numpy(), rtol=1e-12)
    blas.check("blasdsp_stream_release", blas.lib().blasdsp_stream_release(s.cuda_stream))
    # releasing again (nothing left) is fine
    blas.check("blasdsp_stream_release", blas.lib().blasdsp_stream_release(s.cuda_stream))
    C.zero_()
    with torch.cuda.stream(s):
        blas.gemm("N", "N", A, B, C, 1.0, 0.0, algo=blas.ALGO_TILE, splitk=16, stream=s.cuda_stream)
    s.synchronize()
    assert np.allclose(C.cpu().numpy(), ref, rtol=1e-12)
    blas.check("blasdsp_stream_release", blas.lib().blasdsp_stream_release(s.cuda_stream))
