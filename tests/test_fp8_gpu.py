"""Scaled-FP8 FP32 GEMM (blasdsp_sgemm_fp8) against the oracle's restatement
of the same quantization (oracle.gemm_fp8), and its error against FP64.

Exact checks pin the quantizer, the scales and the operand layout: one
product per output reproduces every E4M3 code of the oracle's encoder, and
inputs whose products in each group of 32 k stay within a factor 2^13 of each
other give results bit-identical to the oracle.

The matrix cores sum each group of 32 FP8 products (one lane group's k) in a
fixed-point format whose quantum is 2^-13 of the group's largest product,
truncating; groups are then added in FP32 (measured: 448 + s keeps s only to
multiples of 2^-5, while 448 + s split across two groups is exact).  On
general inputs the library therefore differs from the oracle by up to
32 2^-13 of each group's largest product: bar 2^-11 normwise (measured
4.9e-6 .. 7.3e-6 on U[-1,1)).  Against the unquantized FP64 product the error
is E4M3's; it is reported and must lie in the band E4M3 rounding implies."""
import math

import numpy as np
import pytest

from conftest import to_dev

pytestmark = pytest.mark.gpu


def tol32(K):
    return (math.sqrt(max(K, 1)) + 4) * 2.0 ** -23


# group sums of the FP8 matrix cores: quantum 2^-13 of the group's largest product
TOL_GROUP = 2.0 ** -11


def run(orc, blas, dev, mode, M, N, K, dist=0, alpha=1.5, beta=-0.5, seed=51):
    import torch

    A = orc.matrix(M, K, dist, seed, dtype=np.float32, ld=(M if mode[0] == "T" else K) + 3, trans=mode[0] == "T")
    B = orc.matrix(K, N, dist, seed + 1, dtype=np.float32, ld=(K if mode[1] == "T" else N) + 1, trans=mode[1] == "T")
    C0 = orc.matrix(M, N, 0, seed + 2, dtype=np.float32, ld=N + 2)
    ref = orc.gemm_fp8(mode[0], mode[1], A, B, C0, alpha, beta)
    dC = to_dev(C0, dev)
    blas.sgemm_fp8(mode[0], mode[1], to_dev(A, dev), to_dev(B, dev), dC, alpha, beta)
    torch.cuda.synchronize()
    return np.ascontiguousarray(dC.cpu().numpy()), ref, (A, B, C0)


@pytest.mark.parametrize("mode", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("shape", [(1, 1, 1), (33, 17, 70), (128, 128, 128), (200, 136, 300), (257, 129, 1000),
                                   (64, 300, 4097)])
def test_fp8_matches_quantized_oracle(orc, blas, dev, mode, shape):
    got, ref, _ = run(orc, blas, dev, mode, *shape)
    err, _ = orc.normwise(got, ref)
    assert err <= TOL_GROUP, err


@pytest.mark.parametrize("mode", ["NN", "NT", "TN", "TT"])
def test_fp8_bit_exact_without_group_truncation(orc, blas, dev, mode):
    """Magnitudes in [0.5, 1) times a power of two per operand, random signs:
    every product in a group is within 2^2 of the largest, nothing is
    truncated, and the result equals the oracle bit for bit (any quantizer,
    scale or layout mistake shows)."""
    import torch

    M, N, K = 200, 136, 1000
    rng = np.random.default_rng(71)

    def mat(rows, cols, scale):
        x = rng.uniform(0.5, 1.0, (rows, cols)) * rng.choice([-1.0, 1.0], (rows, cols)) * scale
        return x.astype(np.float32)

    A = mat(M, K, 2.0 ** -7)
    B = mat(K, N, 2.0 ** 9)
    Ast = np.ascontiguousarray(A.T) if mode[0] == "T" else A
    Bst = np.ascontiguousarray(B.T) if mode[1] == "T" else B
    C0 = np.zeros((M, N), dtype=np.float32)
    ref = orc.gemm_fp8(mode[0], mode[1], Ast, Bst, C0, 1.0, 0.0)
    dC = to_dev(C0, dev)
    blas.sgemm_fp8(mode[0], mode[1], to_dev(Ast, dev), to_dev(Bst, dev), dC, 1.0, 0.0)
    torch.cuda.synchronize()
    # every product has at most 8 significant bits within a 2^4 exponent
    # range, so all 1000-term sums are exact in FP32 as in the oracle
    np.testing.assert_array_equal(dC.cpu().numpy().astype(np.float64), ref)


def test_fp8_quantizer_every_magnitude(orc, blas, dev):
    """One product per output: C[i] = E4M3(v_i) exactly, for 2000 magnitudes
    over the whole E4M3 range (a fixed 448 in the block pins the scale to 1)."""
    import torch

    rng = np.random.default_rng(72)
    v = np.concatenate([np.exp2(rng.uniform(-12, np.log2(440), 2000)), np.exp2(np.arange(-12, 9, 0.25))])
    v = (v * rng.choice([-1, 1], v.size)).astype(np.float32)
    M = v.size
    A = np.zeros((M, 128), np.float32)
    A[:, 0] = v
    A[:, 127] = 448.0
    B = np.zeros((128, 1), np.float32)
    B[0, 0] = 1.0
    C = torch.zeros((M, 1), dtype=torch.float32, device=dev)
    blas.sgemm_fp8("N", "N", to_dev(A, dev), to_dev(B, dev), C, 1.0, 0.0)
    torch.cuda.synchronize()
    want = np.array([orc.fp8_decode(orc.fp8_e4m3(float(x))) for x in v], np.float32)
    np.testing.assert_array_equal(C.cpu().numpy()[:, 0], want)


def test_fp8_quantizer_rounding_boundaries(orc, blas, dev):
    """Every tie between adjacent E4M3 codes, and the FP32 values on either
    side of it, of both signs and through the subnormals: the quantizer rounds
    each to nearest, ties to even, as the oracle's encoder does."""
    import torch

    codes = np.arange(0, 0x7F)  # 0 .. 448, positive finite codes
    vals = np.array([orc.fp8_decode(int(q)) for q in codes], np.float64)
    mids = ((vals[:-1] + vals[1:]) / 2).astype(np.float32)  # exact in FP32
    v = np.concatenate([mids, np.nextafter(mids, np.float32(0)), np.nextafter(mids, np.float32(np.inf)),
                        vals.astype(np.float32)])
    v = np.concatenate([v, -v]).astype(np.float32)
    M = v.size
    A = np.zeros((M, 128), np.float32)
    A[:, 0] = v
    A[:, 127] = 448.0
    B = np.zeros((128, 1), np.float32)
    B[0, 0] = 1.0
    C = torch.zeros((M, 1), dtype=torch.float32, device=dev)
    blas.sgemm_fp8("N", "N", to_dev(A, dev), to_dev(B, dev), C, 1.0, 0.0)
    torch.cuda.synchronize()
    want = np.array([orc.fp8_decode(orc.fp8_e4m3(float(x))) for x in v], np.float32)
    np.testing.assert_array_equal(C.cpu().numpy()[:, 0], want)


def test_fp8_group_sum_quantum(orc, blas, dev):
    """The property the general bar rests on: within one group of 32 k the
    sum 448 + s loses at most 2^-5 = 2^-13 * 2^8; across two groups it is
    exact."""
    import torch

    s = np.array([orc.fp8_decode(c) for c in range(1, 0x7F)], np.float32)
    M = s.size
    for k_small, exact in ((1, False), (64, True)):
        A = np.zeros((M, 128), np.float32)
        A[:, 0] = 448.0
        A[:, k_small] = s
        B = np.zeros((128, 1), np.float32)
        B[0, 0] = B[k_small, 0] = 1.0
        C = torch.zeros((M, 1), dtype=torch.float32, device=dev)
        blas.sgemm_fp8("N", "N", to_dev(A, dev), to_dev(B, dev), C, 1.0, 0.0)
        torch.cuda.synchronize()
        got = C.cpu().numpy()[:, 0].astype(np.float64)
        want = 448.0 + s.astype(np.float64)
        if exact:
            np.testing.assert_array_equal(got, want)
        else:
            assert np.all(np.abs(got - want) < 2.0 ** -5)


@pytest.mark.parametrize("dist", [0, 1, 2])
def test_fp8_error_vs_fp64(orc, blas, dev, dist):
    """Uniform, normal and log-uniform (2^[-8,8]) inputs, K = 2048."""
    got, ref, (A, B, C0) = run(orc, blas, dev, "NN", 96, 80, 2048, dist=dist, alpha=1.0, beta=0.0)
    exact = orc.gemm_xp("N", "N", A, B, C0, 1.0, 0.0)
    err, _ = orc.normwise(got, exact)
    print(f"fp8 normwise error vs FP64, dist {dist}: {err:.3e}")
    assert 0.002 < err < 0.08, err
    qerr, _ = orc.normwise(got, ref)
    assert qerr <= TOL_GROUP, qerr


def test_fp8_integer_patterns_exact(orc, blas, dev):
    """%11 x %7 patterns survive E4M3 exactly and every sum stays below 2^24."""
    import torch

    M, N, K = 130, 140, 1500
    P = (orc.matrix(M, K, orc.PATTERN11, 0, dtype=np.float32).astype(np.int64)
         @ orc.matrix(K, N, orc.PATTERN7, 0, dtype=np.float32).astype(np.int64))
    for mode in ("NN", "NT", "TN", "TT"):
        A = orc.matrix(M, K, orc.PATTERN11, 0, dtype=np.float32, trans=mode[0] == "T")
        B = orc.matrix(K, N, orc.PATTERN7, 0, dtype=np.float32, trans=mode[1] == "T")
        C = torch.full((M, N), float("nan"), dtype=torch.float32, device=dev)
        blas.sgemm_fp8(mode[0], mode[1], to_dev(A, dev), to_dev(B, dev), C, 1.0, 0.0)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(C.cpu().numpy(), P.astype(np.float32), err_msg=mode)


def test_fp8_per_block_scales_wide_range(orc, blas, dev):
    """Rows 2^40 apart and K blocks 2^40 apart: every block keeps its own
    scale, so the library still matches the quantized oracle."""
    import torch

    M, N, K = 40, 50, 512
    A = orc.matrix(M, K, 0, 61, dtype=np.float32)
    B = orc.matrix(K, N, 0, 62, dtype=np.float32)
    A = (A * np.float32(2.0) ** np.linspace(-20, 20, M, dtype=np.float32)[:, None]).astype(np.float32)
    A[:, 256:] *= np.float32(2.0 ** 20)
    A[:, :128] *= np.float32(2.0 ** -20)
    B = (B * np.float32(2.0) ** np.linspace(20, -20, N, dtype=np.float32)[None, :]).astype(np.float32)
    C0 = np.zeros((M, N), dtype=np.float32)
    ref = orc.gemm_fp8("N", "N", A, B, C0, 1.0, 0.0)
    dC = to_dev(C0, dev)
    blas.sgemm_fp8("N", "N", to_dev(A, dev), to_dev(B, dev), dC, 1.0, 0.0)
    torch.cuda.synchronize()
    got = dC.cpu().numpy().astype(np.float64)
    rel = np.abs(got - ref) / np.abs(ref).max(axis=1, keepdims=True)
    assert rel.max() <= TOL_GROUP, rel.max()


@pytest.mark.parametrize("mode", ["NN", "NT", "TN", "TT"])
def test_fp8_wide_tiles_match_small_launches(orc, blas, dev, mode):
    """M = 8292, N = 8000, K = 300: at least four 256 x 256 tiles per CU, so
    the 256-tile kernel runs; each 100-row band (top, middle, and the ragged
    last tile row) equals, bit for bit, the same rows run as a 100-row problem
    (the 128-tile kernel), and matches the quantized oracle."""
    import torch

    M, N, K, nb = 8292, 8000, 300, 100
    ta, tb = mode[0] == "T", mode[1] == "T"
    A = torch.empty((K, M) if ta else (M, K), dtype=torch.float32, device=dev)
    B = torch.empty((N, K) if tb else (K, N), dtype=torch.float32, device=dev)
    C = torch.empty((M, N), dtype=torch.float32, device=dev)
    blas.fill(A, M, K, blas.LOGUNIF, 81, trans=ta)
    blas.fill(B, K, N, blas.LOGUNIF, 82, trans=tb)
    blas.fill(C, M, N, blas.UNIFORM, 83)
    C0 = C.clone()
    blas.sgemm_fp8(mode[0], mode[1], A, B, C, 1.5, -0.5)
    hB = np.ascontiguousarray(B.cpu().numpy())
    for r0 in (0, 4100, M - nb):
        Ab = A[:, r0:r0 + nb] if ta else A[r0:r0 + nb]
        Cb = C0[r0:r0 + nb].clone()
        blas.sgemm_fp8(mode[0], mode[1], Ab, B, Cb, 1.5, -0.5)
        torch.cuda.synchronize()
        got = np.ascontiguousarray(C[r0:r0 + nb].cpu().numpy())
        np.testing.assert_array_equal(got, Cb.cpu().numpy(), err_msg=f"{mode} rows {r0}")
        ref = orc.gemm_fp8(mode[0], mode[1], np.ascontiguousarray(Ab.cpu().numpy()), hB,
                           np.ascontiguousarray(C0[r0:r0 + nb].cpu().numpy()), 1.5, -0.5)
        err, _ = orc.normwise(got, ref)
        assert err <= TOL_GROUP, (mode, r0, err)


def test_fp8_alpha_zero_and_empty(orc, blas, dev):
    import torch

    C = torch.ones((5, 7), dtype=torch.float32, device=dev)
    A = torch.full((5, 3), float("nan"), dtype=torch.float32, device=dev)
    B = torch.full((3, 7), float("nan"), dtype=torch.float32, device=dev)
    blas.sgemm_fp8("N", "N", A, B, C, 0.0, 2.0)
    torch.cuda.synchronize()
    assert torch.equal(C.cpu(), torch.full((5, 7), 2.0))
    E = torch.empty((0, 7), dtype=torch.float32, device=dev)
    blas.sgemm_fp8("N", "N", torch.empty((0, 3), device=dev), B, E, 1.0, 0.0)
