"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Each check is independent of the oracle's own formulas: exact int64 matrix
products of the reference demos' integer patterns, hand-derived rounding cases
that fix the chunk structure of the reference host loops, exact rational
brute force, closed forms and numpy's own norms.
"""
from fractions import Fraction

import numpy as np
import pytest


def exact_int(A, B):
    return np.asarray(A, dtype=np.int64) @ np.asarray(B, dtype=np.int64)


# ---- generators --------------------------------------------------------------

def test_generator_logical_index_and_padding(orc):
    A = orc.matrix(37, 23, orc.UNIFORM, 7, ld=40)
    AT = orc.matrix(37, 23, orc.UNIFORM, 7, ld=50, trans=True)
    assert A.shape == (37, 23) and AT.shape == (23, 37)
    np.testing.assert_array_equal(A, AT.T)
    assert np.isnan(A.base[:, 23:]).all() and np.isnan(AT.base[:, 37:]).all()
    assert (A >= -1).all() and (A < 1).all()
    A32 = orc.matrix(37, 23, orc.UNIFORM, 7, dtype=np.float32)
    np.testing.assert_array_equal(A32, A.astype(np.float32))


def test_generator_distributions(orc):
    u = orc.matrix(512, 512, orc.UNIFORM, 0)
    assert abs(u.mean()) < 0.01 and abs(u.var() - 1 / 3) < 0.01
    z = orc.matrix(512, 512, orc.NORMAL, 2)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    g = orc.matrix(512, 512, orc.LOGUNIF, 3)
    e = np.log2(np.abs(g))
    assert e.min() >= -8 and e.max() < 8 and abs(e.mean()) < 0.05
    assert abs((g > 0).mean() - 0.5) < 0.01
    # different seeds differ, same seed repeats
    assert not np.array_equal(orc.matrix(8, 8, orc.UNIFORM, 0), orc.matrix(8, 8, orc.UNIFORM, 1))
    np.testing.assert_array_equal(orc.matrix(8, 8, orc.UNIFORM, 5), orc.matrix(8, 8, orc.UNIFORM, 5))


def test_generator_patterns(orc):
    A = orc.matrix(5, 13, orc.PATTERN11, 0)
    i, j = np.meshgrid(np.arange(5), np.arange(13), indexing="ij")
    np.testing.assert_array_equal(A, (i * 13 + j) % 11)


# ---- known answers: the reference demos' integer matrices --------------------
# A = (i*K + k) % 11, B = (k*N + j) % 7
# (irr_shortA_dgemm_release_V0.0/host_code/irr_shortA_dgemm_demo.c:67,76,
#  irr_thinB_sgemm_release_V1.2/host_code/irr_thinB_sgemm_demo.c:77,86)

@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (6, 64, 512), (13, 29, 1100), (32, 48, 3072 + 700)])
def test_kat_integer_patterns(orc, M, N, K):
    A = orc.matrix(M, K, orc.PATTERN11, 0, ld=K + 3)
    B = orc.matrix(K, N, orc.PATTERN7, 0, ld=N + 5)
    P = exact_int(A, B)

    C = np.zeros((M, N))
    orc.dgemm_v11_host(A, B, C)
    np.testing.assert_array_equal(C, P)

    C = np.zeros((M, N), dtype=np.float32)
    orc.sgemm_plain_host(A.astype(np.float32), B.astype(np.float32), C)
    np.testing.assert_array_equal(C, P)

    # sgemm V1.3.2 with the launcher's alpha = 5, beta = 3
    # (sgemm_release_V1.3.2/device_code/sgemm_launcher.c:11-12)
    C = np.zeros((M, N), dtype=np.float32)
    orc.sgemm_v132_host(A.astype(np.float32), B.astype(np.float32), C, 5.0, 3.0)
    np.testing.assert_array_equal(C, 5 * P + 3)

    C = np.zeros((M, N), dtype=np.float32)
    orc.thinB_sgemm_v12_host(A.astype(np.float32), B.astype(np.float32), C, 5.0, 3.0)
    np.testing.assert_array_equal(C, 5 * P + 3)

    C0 = orc.matrix(M, N, orc.PATTERN7, 9)
    C = C0.copy()
    orc.gemm_blas_host(A, B, C, 1.5, -0.5)
    np.testing.assert_array_equal(C, 1.5 * P - 0.5 * C0)
    C = C0.astype(np.float32)
    orc.gemm_blas_host(A.astype(np.float32), B.astype(np.float32), C, 1.5, -0.5)
    np.testing.assert_array_equal(C, 1.5 * P - 0.5 * C0)

    for ta in "NT":
        for tb in "NT":
            Aop = orc.matrix(M, K, orc.PATTERN11, 0, trans=ta == "T", ld=(M if ta == "T" else K) + 1)
            Bop = orc.matrix(K, N, orc.PATTERN7, 0, trans=tb == "T", ld=(K if tb == "T" else N) + 2)
            R = orc.gemm_xp(ta, tb, Aop, Bop, C0, 2.0, 3.0)
            np.testing.assert_array_equal(R, 2 * P + 3 * C0)
            R = orc.gemm_xp(ta, tb, Aop.astype(np.float32), Bop.astype(np.float32),
                            C0.astype(np.float32), 2.0, 3.0)
            np.testing.assert_array_equal(R, 2 * P + 3 * C0)


# ---- rounding cases that fix the reference's accumulation tree ---------------

def _row_case(K, head, dtype):
    """A = 1 x K row [head, 1, 1, ...], B = K x 1 ones."""
    A = np.ones((1, K), dtype=dtype)
    A[0, 0] = head
    B = np.ones((K, 1), dtype=dtype)
    return A, B


def test_chunks_dgemm_v11(orc):
    # k = 0 .. 511 is one chunk: 2^53 + 1 + ... rounds back to 2^53 at every
    # step (ties-to-even); the tail chunk k = 512, 513 sums to 2 exactly, so
    # the result is 2^53 + 2.  A chain without chunks gives 2^53; K_G = 256
    # gives 2^53 + 256 + 2; dropping the tail gives 2^53.
    A, B = _row_case(514, 2.0 ** 53, np.float64)
    C = np.zeros((1, 1))
    orc.dgemm_v11_host(A, B, C)
    assert C[0, 0] == 2.0 ** 53 + 2
    C = np.zeros((1, 1))
    orc.dgemm_v11_host(A, B, C, K_G=256)
    assert C[0, 0] == 2.0 ** 53 + 256 + 2
    # the BLAS form keeps the same tree
    C = np.zeros((1, 1))
    orc.gemm_blas_host(A, B, C, 1.0, 0.0)
    assert C[0, 0] == 2.0 ** 53 + 2


def test_chunks_sgemm(orc):
    # float tmp: fma runs in double, the assignment rounds to float, so
    # 2^24 + 1 -> 2^24 inside a chunk; the tail adds 2.
    A, B = _row_case(514, 2.0 ** 24, np.float32)
    for fn in (orc.sgemm_plain_host,):
        C = np.zeros((1, 1), dtype=np.float32)
        fn(A, B, C)
        assert C[0, 0] == 2.0 ** 24 + 2
    C = np.zeros((1, 1), dtype=np.float32)
    orc.sgemm_v132_host(A, B, C, 1.0, 0.0)
    assert C[0, 0] == 2.0 ** 24 + 2
    C = np.zeros((1, 1), dtype=np.float32)
    orc.gemm_blas_host(A, B, C, 1.0, 0.0)
    assert C[0, 0] == 2.0 ** 24 + 2


def test_chunks_thinB_v12(orc):
    # K_G = 3072 outer, K_A = 512 inner chunks.  head 2^24 then 3071 ones:
    # inner chunk 0 (k < 512) = 2^24, inner chunks 1..5 = 512 each (exact);
    # tmp_i adds them in order: 2^24 + 512 * 5 exactly.  The next outer chunk
    # k = 3072 .. 3073 sums to 2; tmp_o = 2^24 + 2560 + 2.
    A, B = _row_case(3074, 2.0 ** 24, np.float32)
    C = np.zeros((1, 1), dtype=np.float32)
    orc.thinB_sgemm_v12_host(A, B, C, 1.0, 0.0)
    assert C[0, 0] == 2.0 ** 24 + 2560 + 2
    C = np.zeros((1, 1), dtype=np.float32)
    orc.thinB_sgemm_v12_host(A, B, C, 1.0, 0.0, K_G=3072, K_A=3072)
    assert C[0, 0] == 2.0 ** 24 + 2  # one inner chunk: all 3071 ones lost


def _f32(x):
    return Fraction(float(np.float32(float(x))))


def test_alpha_folding_sites(orc):
    # sgemm V1.3.2 rounds alpha*A to float before the fma
    # (sgemm_demo.c:228); thinB V1.2 rounds alpha*B (irr_thinB_sgemm_demo.c:231).
    # With a = 1 + 2^-23, b = 1 + 2 * 2^-23, alpha = 5 the two sites give
    # different floats; the expected values below follow from exact rational
    # arithmetic with one rounding per C operation (the products of floats
    # are exact in double, so the double fma adds no rounding of its own).
    a, b, alpha = 1 + 2.0 ** -23, 1 + 2.0 ** -22, 5.0
    A = np.array([[a]], dtype=np.float32)
    B = np.array([[b]], dtype=np.float32)
    fold_a = _f32(_f32(Fraction(alpha) * Fraction(a)) * Fraction(b))
    fold_b = _f32(Fraction(a) * _f32(Fraction(alpha) * Fraction(b)))
    assert fold_a != fold_b  # the case discriminates
    C = np.zeros((1, 1), dtype=np.float32)
    orc.sgemm_v132_host(A, B, C, alpha, 0.0)
    assert Fraction(float(C[0, 0])) == fold_a
    C = np.zeros((1, 1), dtype=np.float32)
    orc.thinB_sgemm_v12_host(A, B, C, alpha, 0.0)
    assert Fraction(float(C[0, 0])) == fold_b


def test_scalar_beta_vs_blas_beta(orc):
    # reference: C_cpu starts at 0 (sgemm_demo.c:87) and gets + beta; the BLAS
    # form scales the input C.
    A = np.ones((2, 3), dtype=np.float32)
    B = np.ones((3, 2), dtype=np.float32)
    C = np.full((2, 2), 10.0, dtype=np.float32)
    orc.sgemm_v132_host(A, B, C, 2.0, 3.0)
    np.testing.assert_array_equal(C, 10 + 6 + 3)
    C = np.full((2, 2), 10.0, dtype=np.float32)
    orc.gemm_blas_host(A, B, C, 2.0, 3.0)
    np.testing.assert_array_equal(C, 6 + 30)


# ---- extended-precision oracle vs exact rationals ----------------------------

@pytest.mark.parametrize("ta", "NT")
@pytest.mark.parametrize("tb", "NT")
def test_xp_vs_fractions(orc, ta, tb):
    rng = np.random.default_rng(11)
    M, N, K = 5, 4, 7
    alpha, beta = 0.7, -1.3
    Aop = rng.standard_normal((M, K))
    Bop = rng.standard_normal((K, N))
    C0 = rng.standard_normal((M, N))
    A = np.ascontiguousarray(Aop.T) if ta == "T" else Aop
    B = np.ascontiguousarray(Bop.T) if tb == "T" else Bop
    R = orc.gemm_xp(ta, tb, A, B, C0, alpha, beta)
    for i in range(M):
        for j in range(N):
            s = sum(Fraction(Aop[i, k]) * Fraction(Bop[k, j]) for k in range(K))
            exact = Fraction(alpha) * s + Fraction(beta) * Fraction(C0[i, j])
            ex = float(exact)
            assert abs(R[i, j] - ex) <= np.spacing(abs(ex)), (i, j)


def test_xp_f32_inputs_exact_products(orc):
    rng = np.random.default_rng(3)
    A = rng.standard_normal((6, 9)).astype(np.float32)
    B = rng.standard_normal((9, 5)).astype(np.float32)
    C = rng.standard_normal((6, 5)).astype(np.float32)
    R = orc.gemm_xp("N", "N", A, B, C, 1.0, 1.0)
    for i in range(6):
        for j in range(5):
            s = sum(Fraction(float(A[i, k])) * Fraction(float(B[k, j])) for k in range(9))
            ex = float(s + Fraction(float(C[i, j])))
            assert abs(R[i, j] - ex) <= np.spacing(abs(ex))


# ---- closed forms -------------------------------------------------------------

def test_closed_forms(orc):
    M, N, K = 9, 11, 37
    C0 = orc.matrix(M, N, orc.UNIFORM, 4)
    ones_a = orc.matrix(M, K, orc.ONES, 0)
    ones_b = orc.matrix(K, N, orc.ONES, 0)
    np.testing.assert_array_equal(orc.gemm_xp("N", "N", ones_a, ones_b, C0, 2.0, 0.5), 2.0 * K + 0.5 * C0)
    A = orc.matrix(M, K, orc.UNIFORM, 5)
    eye = np.eye(K)
    np.testing.assert_array_equal(orc.gemm_xp("N", "N", A, eye, np.zeros((M, K)), 1.0, 0.0), A)
    # K = 0: C = beta C
    np.testing.assert_array_equal(orc.gemm_xp("N", "N", np.zeros((M, 0)), np.zeros((0, N)), C0, 1.0, 2.0),
                                  2.0 * C0)
    # rank one: u 1^T times 1 v^T = K u v^T
    u = np.arange(1, M + 1, dtype=np.float64)[:, None]
    v = np.arange(1, N + 1, dtype=np.float64)[None, :]
    R = orc.gemm_xp("N", "N", u * np.ones((1, K)), np.ones((K, 1)) * v, np.zeros((M, N)), 1.0, 0.0)
    np.testing.assert_array_equal(R, K * u * v)


def test_normwise_matches_numpy(orc):
    rng = np.random.default_rng(0)
    R = rng.standard_normal((30, 20))
    X = R + 1e-7 * rng.standard_normal((30, 20))
    err, mx = orc.normwise(X, R)
    assert err == pytest.approx(np.linalg.norm(X - R) / np.linalg.norm(R), rel=1e-12)
    assert mx == pytest.approx(np.abs(X - R).max(), rel=1e-12)
    err32, _ = orc.normwise(X.astype(np.float32), R)
    assert err32 == pytest.approx(np.linalg.norm(X.astype(np.float32) - R) / np.linalg.norm(R), rel=1e-9)


# ---- random projection (whole-result checker) ----------------------------------

@pytest.mark.parametrize("trans", "NT")
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_project_vs_fractions(orc, trans, dt):
    rng = np.random.default_rng(21)
    M, K, S = 5, 7, 3
    Aop = rng.standard_normal((M, K)).astype(dt)
    X = rng.standard_normal((K, S))
    A = np.ascontiguousarray(Aop.T) if trans == "T" else Aop
    Y = orc.project(trans, A, X)
    for i in range(M):
        for s in range(S):
            ex = float(sum(Fraction(float(Aop[i, k])) * Fraction(X[k, s]) for k in range(K)))
            assert abs(Y[i, s] - ex) <= np.spacing(abs(ex)), (i, s)


@pytest.mark.parametrize("trans", "NT")
def test_project_padded_blocks_and_limits(orc, trans):
    """Row padding, a column count that leaves a partial block of the
    transposed walk (70 = 2 x 32 + 6), every S from 1 to 16, and S = 17
    refused."""
    rng = np.random.default_rng(22)
    M, K = 70, 45
    Aop = rng.standard_normal((M, K))
    store = np.zeros((K, M + 5) if trans == "T" else (M, K + 3))
    A = store[:, :M] if trans == "T" else store[:, :K]
    A[...] = Aop.T if trans == "T" else Aop
    for S in (1, 7, 16):
        X = rng.standard_normal((K, S))
        np.testing.assert_allclose(orc.project(trans, A, X), Aop @ X, rtol=1e-12, atol=1e-13)
    with pytest.raises(AssertionError):
        orc.project(trans, A, rng.standard_normal((K, 17)))


def _projection_case(orc, M, N, K, seed=23, S=8):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    C0 = rng.standard_normal((M, N))
    R = orc.gemm_xp("N", "N", A, B, C0, 1.5, -0.5)
    X = rng.standard_normal((N, S))
    RX = orc.gemm_projection("N", A, "N", B, C0, 1.5, -0.5, X)
    return A, B, C0, R, X, RX, rng


def test_projected_errors_estimate_the_normwise_error(orc):
    """Planted errors of known size: the estimate matches the directly
    measured normwise error; an exact result reads at rounding level."""
    A, B, C0, R, X, RX, rng = _projection_case(orc, 300, 200, 150)
    nw, row, _ = orc.projected_errors(R, RX, X)
    assert nw < 1e-15 and row < 1e-14, (nw, row)
    G = R + 1e-8 * np.abs(R).mean() * rng.standard_normal(R.shape)
    true, _ = orc.normwise(G, R)
    nw, _, _ = orc.projected_errors(G, RX, X)
    assert abs(nw / true - 1) < 0.1, (nw, true)
    # one k term dropped from every output: a rank-one error, whose projection
    # has only S degrees of freedom (estimate = true x chi_S / sqrt(S))
    G = 1.5 * (A[:, 1:] @ B[1:, :]) - 0.5 * C0
    true, _ = orc.normwise(G, R)
    nw, _, _ = orc.projected_errors(G, RX, X)
    assert true > 1e-3 and 0.4 < nw / true < 1.8, (nw, true)
    # op(A) transposed by mistake (square case)
    A2, B2, C2, R2, X2, RX2, _ = _projection_case(orc, 120, 90, 120, seed=24)
    nw, _, _ = orc.projected_errors(1.5 * (A2.T @ B2) - 0.5 * C2, RX2, X2)
    assert nw > 0.3, nw


def test_projected_errors_find_one_wrong_element(orc):
    """A single output off by 1e-9 of its row's norm is found in its row by the
    right projection and in its column by the left one, while every other
    row reads at rounding level."""
    A, B, C0, R, X, RX, rng = _projection_case(orc, 300, 200, 150)
    i0, j0 = 217, 41
    G = R.copy()
    G[i0, j0] += 1e-9 * np.linalg.norm(R[i0])
    nw, row, i = orc.projected_errors(G, RX, X)
    assert i == i0 and row > 2e-10, (i, row)
    E = orc.project("N", G, X) - RX
    mask = np.arange(R.shape[0]) != i0
    others = np.linalg.norm(E[mask], axis=1) / (np.sqrt(X.shape[1]) * np.linalg.norm(G[mask], axis=1))
    assert others.max() < 1e-14, others.max()
    Y = rng.standard_normal((R.shape[0], 8))
    RY = orc.gemm_projection("N", A, "N", B, C0, 1.5, -0.5, Y, side="left")
    np.testing.assert_allclose(RY, R.T @ Y, rtol=1e-12, atol=1e-10)
    nw_l, col, j = orc.projected_errors(R, RY, Y, side="left")
    assert nw_l < 1e-15 and col < 1e-14, (nw_l, col)
    G[i0, j0] = R[i0, j0] + 1e-9 * np.linalg.norm(R[:, j0])
    _, col, j = orc.projected_errors(G, RY, Y, side="left")
    assert j == j0 and col > 2e-10, (j, col)


def test_projected_errors_fail_on_nan_and_inf(orc):
    """A NaN or an infinity anywhere in the result makes both estimates
    non-finite, so every bar written as `estimate <= bar` fails."""
    A, B, C0, R, X, RX, _ = _projection_case(orc, 60, 50, 40)
    for bad in (np.nan, np.inf):
        G = R.copy()
        G[37, 11] = bad
        nw, row, i = orc.projected_errors(G, RX, X)
        assert not nw <= 1.0 and not row <= 1.0, (bad, nw, row)


def test_gemm_projection_modes_and_f32(orc):
    """Every storage mode of A and B gives the same R X; FP32 operands
    project exactly as their FP64 values do."""
    rng = np.random.default_rng(25)
    M, N, K, S = 40, 33, 57, 4
    Aop = rng.standard_normal((M, K)).astype(np.float32)
    Bop = rng.standard_normal((K, N)).astype(np.float32)
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    X = rng.standard_normal((N, S))
    R = orc.gemm_xp("N", "N", Aop, Bop, C0, 2.0, 0.25)
    want = R @ X
    base = orc.gemm_projection("N", Aop.astype(np.float64), "N", Bop.astype(np.float64), C0.astype(np.float64),
                               2.0, 0.25, X)
    Y = rng.standard_normal((M, S))
    base_l = orc.gemm_projection("N", Aop, "N", Bop, C0, 2.0, 0.25, Y, side="left")
    want_l = R.T @ Y
    for ta in "NT":
        for tb in "NT":
            A = np.ascontiguousarray(Aop.T) if ta == "T" else Aop
            B = np.ascontiguousarray(Bop.T) if tb == "T" else Bop
            RX = orc.gemm_projection(ta, A, tb, B, C0, 2.0, 0.25, X)
            np.testing.assert_array_equal(RX, base)
            assert np.linalg.norm(RX - want) / np.linalg.norm(want) < 1e-14
            RY = orc.gemm_projection(ta, A, tb, B, C0, 2.0, 0.25, Y, side="left")
            np.testing.assert_array_equal(RY, base_l)
            assert np.linalg.norm(RY - want_l) / np.linalg.norm(want_l) < 1e-14


def test_mode_invariance_with_padding(orc):
    # the reference harness builds explicit transposed copies and checks all
    # four modes against one NN result (dgemm_release_V1.1/host_code/dgemm_demo.c:74,82,101-131)
    M, N, K = 33, 17, 45
    C0 = orc.matrix(M, N, orc.UNIFORM, 3)
    ref = None
    for ta in "NT":
        for tb in "NT":
            A = orc.matrix(M, K, orc.UNIFORM, 1, trans=ta == "T", ld=64)
            B = orc.matrix(K, N, orc.UNIFORM, 2, trans=tb == "T", ld=64)
            R = orc.gemm_xp(ta, tb, A, B, C0, 1.5, -0.5)
            assert np.isfinite(R).all()
            if ref is None:
                ref = R
            np.testing.assert_array_equal(R, ref)


# ---- scaled FP8 restatement ------------------------------------------------

def _fp8_sample(rng, n):
    """Floats in [-448, 448]: random magnitudes over the whole E4M3 range, exact
    ties between neighbouring codes, subnormals and the extremes."""
    import torch

    mags = np.exp2(rng.uniform(-12, np.log2(448.0), n)).astype(np.float32)
    vals = mags * rng.choice([-1.0, 1.0], n).astype(np.float32)
    codes = torch.arange(0, 127, dtype=torch.uint8).view(torch.float8_e4m3fn).float().numpy()
    ties = ((codes[:-1] + codes[1:]) / 2).astype(np.float32)  # midpoints are exact in float32
    extra = np.array([0.0, -0.0, 448.0, -448.0, 2.0 ** -9, 2.0 ** -10, 3 * 2.0 ** -10, 2.0 ** -6], dtype=np.float32)
    return np.concatenate([vals, ties, -ties, codes, extra]).astype(np.float32)


def test_fp8_encoder_matches_torch_e4m3fn(orc):
    """Round-to-nearest-even E4M3 encoding, against PyTorch's independent
    float8_e4m3fn conversion, including every tie between adjacent codes."""
    import torch

    x = _fp8_sample(np.random.default_rng(5), 20000)
    want = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    got = np.array([orc.fp8_e4m3(float(v)) for v in x], dtype=np.uint8)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(float(x[i]), int(got[i]), int(want[i])) for i in bad[:10]]


def test_fp8_decode_every_code(orc):
    import torch

    codes = torch.arange(0, 256, dtype=torch.int32).to(torch.uint8)
    want = codes.view(torch.float8_e4m3fn).float().numpy()
    for c in range(256):
        got = orc.fp8_decode(c)
        if np.isnan(want[c]):
            assert np.isnan(got), c
        else:
            assert got == want[c], (c, got, want[c])
            if c not in (0x80,):  # -0 encodes back to -0 (0x80) as well
                assert orc.fp8_e4m3(got) == c


def test_fp8_block_exponent_is_the_smallest_fitting_power(orc):
    rng = np.random.default_rng(6)
    amax = np.exp2(rng.uniform(-100, 100, 2000)).astype(np.float32)
    amax = np.concatenate([amax, np.float32([448.0, 448.0 * 2 ** 5, 449.0, 224.0, 1.0, 0.875 * 2 ** 9])])
    for a in amax:
        e = orc.fp8_block_exp(float(a))
        assert float(a) * 2.0 ** -e <= 448.0, (a, e)
        assert float(a) * 2.0 ** -(e - 1) > 448.0, (a, e)
    assert orc.fp8_block_exp(0.0) == 0


def test_fp8_gemm_integer_patterns_exact(orc):
    """The %11 / %7 matrices are E4M3-exact after any power-of-two scaling
    (integers below 16 have at most four significant bits): the quantized
    product is the exact integer product, in every mode."""
    M, N, K = 37, 29, 300
    P = None
    for ta in "NT":
        for tb in "NT":
            A = orc.matrix(M, K, orc.PATTERN11, 0, dtype=np.float32, trans=ta == "T")
            B = orc.matrix(K, N, orc.PATTERN7, 0, dtype=np.float32, trans=tb == "T")
            if P is None:
                An = orc.matrix(M, K, orc.PATTERN11, 0, dtype=np.float32).astype(np.int64)
                Bn = orc.matrix(K, N, orc.PATTERN7, 0, dtype=np.float32).astype(np.int64)
                P = (An @ Bn).astype(np.float64)
            C0 = np.zeros((M, N), dtype=np.float32)
            got = orc.gemm_fp8(ta, tb, A, B, C0, 1.0, 0.0)
            np.testing.assert_array_equal(got, P)


def test_fp8_gemm_per_block_scales(orc):
    """Two K blocks whose magnitudes differ by 2^40: each block keeps its own
    scale, so values that are E4M3 codes times a power of two survive exactly
    and the result equals the exact product (a single scale per row would
    flush the small block to zero)."""
    from fractions import Fraction

    rng = np.random.default_rng(7)
    M, N, K = 3, 4, 256
    codes = np.array([1, 1.125, 1.25, 1.5, 1.75, 2, 3, 5, 7, 10, 14], dtype=np.float32)
    A = rng.choice(codes, (M, K)).astype(np.float32) * rng.choice([-1, 1], (M, K)).astype(np.float32)
    B = rng.choice(codes, (K, N)).astype(np.float32)
    A[:, :128] *= np.float32(2.0 ** -20)
    A[:, 128:] *= np.float32(2.0 ** 20)
    C0 = np.zeros((M, N), dtype=np.float32)
    got = orc.gemm_fp8("N", "N", A, B, C0, 1.0, 0.0)
    for i in range(M):
        for j in range(N):
            exact = sum(Fraction(float(A[i, k])) * Fraction(float(B[k, j])) for k in range(K))
            assert got[i, j] == float(exact), (i, j)


def test_fp8_gemm_error_band(orc):
    """Random normal operands: the quantized product's normwise error against
    the exact product lies in the band E4M3 rounding implies (relative element
    error below 2^-4, about 2^-4/sqrt(3) rms for each operand)."""
    M, N, K = 64, 48, 1000
    A = orc.matrix(M, K, orc.NORMAL, 8, dtype=np.float32)
    B = orc.matrix(K, N, orc.NORMAL, 9, dtype=np.float32)
    C0 = orc.matrix(M, N, orc.NORMAL, 10, dtype=np.float32)
    ref = orc.gemm_xp("N", "N", A, B, C0, 1.0, 0.0)
    got = orc.gemm_fp8("N", "N", A, B, C0, 1.0, 0.0)
    err, _ = orc.normwise(got, ref)
    assert 0.005 < err < 0.08, err


def _dsp_pairing(A):
    """The DSP micro-kernel's pairing of the five terms
    (stencil_2d_V1.2/readme.txt, steps [1]-[18]): rows alternate between
    right + ((left + top) + (middle + down)) and
    right + ((left + down) + (top + middle))."""
    top, left, right = A[:-2, 1:-1], A[1:-1, :-2], A[1:-1, 2:]
    down, mid = A[2:, 1:-1], A[1:-1, 1:-1]
    even = right + ((left + top) + (mid + down))
    odd = right + ((left + down) + (top + mid))
    out = np.where((np.arange(even.shape[0]) % 2 == 0)[:, None], even, odd)
    return out * 0.2


def test_stencil_matches_host_checker_order(orc):
    """The oracle restates reference_stencil (matrix_stencil.c:33-47): against
    numpy's elementwise evaluation of the same left-to-right sum, bit for bit;
    the DSP kernel's pairing differs in the last bits on random data, so the
    order the oracle fixes is the checker's."""
    rng = np.random.default_rng(5)
    for M, N in ((1, 1), (3, 7), (55, 111), (64, 130)):
        A = rng.random((M + 2, N + 2))
        ref = (A[:-2, 1:-1] + A[1:-1, :-2] + A[1:-1, 2:] + A[2:, 1:-1] + A[1:-1, 1:-1]) * 0.2
        np.testing.assert_array_equal(orc.stencil(A), ref)
        H = orc.stencil(A, halo_out=True)
        np.testing.assert_array_equal(H[1:-1, 1:-1], ref)
        np.testing.assert_array_equal(H[0], A[0])
        np.testing.assert_array_equal(H[-1], A[-1])
        np.testing.assert_array_equal(H[:, 0], A[:, 0])
        np.testing.assert_array_equal(H[:, -1], A[:, -1])
    A = rng.random((66, 130))
    assert np.count_nonzero(orc.stencil(A) != _dsp_pairing(A)) > 0


def test_stencil_closed_forms(orc):
    """Constant grid c: every point is c + c + c + c + c (left to right) times
    0.2.  A linear field a*i + b*j is reproduced (the five-point mean of a
    linear function is its centre value) to rounding.  A single unit spike
    spreads 0.2 to itself and its four neighbours only."""
    for c in (1.0, 0.1, -3.75):
        A = np.full((9, 12), c)
        want = (c + c + c + c + c) * 0.2
        assert np.all(orc.stencil(A) == want)
    i, j = np.meshgrid(np.arange(40.0), np.arange(50.0), indexing="ij")
    A = 0.5 * i - 0.25 * j
    np.testing.assert_allclose(orc.stencil(A), A[1:-1, 1:-1], rtol=0, atol=1e-13)
    A = np.zeros((9, 9))
    A[4, 4] = 1.0
    out = orc.stencil(A)
    nz = {(int(r), int(c)) for r, c in zip(*np.nonzero(out))}
    assert nz == {(3, 3), (2, 3), (4, 3), (3, 2), (3, 4)}
    assert np.all(out[out != 0] == 0.2)
