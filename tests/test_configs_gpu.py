"""Parity at the workload configurations' own sizes, through the C ABI,
against the extended-precision oracle.

  c1  SGEMM NN 512^3, lda = K, ldb = ldc = N, alpha = beta = 1, U[-1,1)
      (A, B, C from seeds 0, 1, 2): the whole C, against the oracle and
      against the restated reference host loop (sgemm_demo.c:214-242 with the
      BLAS epilogue); the reference's own semantics (C = alpha AB + beta,
      input C ignored) through sgemm_NN_qxx against the reference-exact
      restatement.
  c3  shortA M in {8, 16, 32, 48}, K = 8192, N = 65536, and thinB M = 65536,
      K = 8192, N in {8, 16, 32, 48}; FP32 and FP64; NN/NT/TN/TT;
      alpha = beta = 1; N(0,1) (long operand seed 2, small seed 3, C seed 4).
      Operands are generated on the device; a 96-wide band of C (columns for
      shortA, rows for thinB) is checked against the oracle on the band's
      operands copied back from the device, and the whole C through the
      oracle's random projection from the long side.
  c4  SGEMM NN 8192^3, U[-1,1) and log-uniform 2^[-8,8] (seeds 3, 4, 5):
      plain FP32, split-BF16 with 6 and 3 terms, split-FP16 and scaled FP8;
      a 32-row band of C against the oracle and the whole C through the
      oracle's random projection, each form with its bar.

Random projection (Freivalds' method, S = 8 Gaussian vectors): C X against
R X = alpha op(A) (op(B) X) + beta C0 X in extended precision, for every
line of C at once; the normwise estimate takes the form's bar, every line's
relative error estimate four times it (FP64: 1e-12).

Bars (normwise ||C - C_ref||_F / ||C_ref||_F):
  FP64 1e-13; FP32 (sqrt(K) + 4) 2^-23; split-BF16 6 terms and split-FP16
  1e-6; split-BF16 3 terms 3e-5; FP8 2^-11 against the oracle's quantized
  product (the matrix cores' 32-product group sums carry 13 bits below the
  group's largest product) and within E4M3's band against FP64.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL64 = 1e-13
LINE64 = 1e-12
PROJ_S = 8


def tol32(K):
    return (math.sqrt(max(K, 1)) + 4) * 2.0 ** -23


def host(t):
    return np.ascontiguousarray(t.cpu().numpy())


# ---- c1 --------------------------------------------------------------------

def test_c1_sgemm_nn_512(orc, blas, dev):
    import torch

    n = 512
    A = orc.matrix(n, n, orc.UNIFORM, 0, dtype=np.float32)
    B = orc.matrix(n, n, orc.UNIFORM, 1, dtype=np.float32)
    C0 = orc.matrix(n, n, orc.UNIFORM, 2, dtype=np.float32)
    dA, dB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    dC = torch.from_numpy(C0).to(dev)
    L = blas.lib()
    s = torch.cuda.current_stream().cuda_stream
    assert L.blasdsp_sgemm_NN(n, n, n, n, n, n, dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), 1.0, 1.0, s) == 0
    torch.cuda.synchronize()
    got = host(dC)
    ref = orc.gemm_xp("N", "N", A, B, C0, 1.0, 1.0)
    err, _ = orc.normwise(got, ref)
    assert err <= tol32(n), err
    # the restated reference host loop (K_G = 512 fma chains in double) with
    # the BLAS epilogue: both within the bar of the exact product
    cpu = C0.copy()
    orc.gemm_blas_host(A, B, cpu, 1.0, 1.0)
    err_cpu, _ = orc.normwise(cpu, ref)
    assert err_cpu <= tol32(n), err_cpu
    diff, _ = orc.normwise(got, cpu.astype(np.float64))
    assert diff <= 2 * tol32(n), diff
    # the reference's own call: C = alpha A B + beta, input C (999) ignored
    dC.fill_(999.0)
    assert L.sgemm_NN_qxx(n, n, n, dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), 1.0, 1.0, s) == 0
    torch.cuda.synchronize()
    lit = np.zeros((n, n), dtype=np.float32)
    orc.sgemm_v132_host(A, B, lit, 1.0, 1.0)
    err_lit, _ = orc.normwise(host(dC), lit.astype(np.float64))
    assert err_lit <= 2 * tol32(n), err_lit


# ---- c3 --------------------------------------------------------------------

K3, LONG = 8192, 65536
BAND = 96


def c3_cases():
    out = []
    for prec in ("s", "d"):
        for kind in ("shortA", "thinB"):
            for mode in ("NN", "NT", "TN", "TT"):
                for small in (8, 16, 32, 48):
                    out.append((prec, kind, mode, small))
    return out


_long_cache = {}


def long_operand(orc, blas, dev, prec, kind, trans):
    """The streamed operand (shortA: op(B), K x LONG; thinB: op(A), LONG x K),
    stored transposed when trans; generated once per (precision, kind,
    orientation) and kept while consecutive cases use it, with its random
    projection from the long side: op(B) X for shortA, op(A)^T X for thinB
    (X: LONG x S, N(0,1), seed 43)."""
    import torch

    key = (prec, kind, trans)
    if key not in _long_cache:
        _long_cache.clear()
        torch.cuda.empty_cache()
        rows, cols = (K3, LONG) if kind == "shortA" else (LONG, K3)
        dt = torch.float64 if prec == "d" else torch.float32
        X = torch.empty((cols, rows) if trans else (rows, cols), dtype=dt, device=dev)
        blas.fill(X, rows, cols, blas.NORMAL, 2, trans=trans)
        P = np.random.default_rng(43).standard_normal((LONG, PROJ_S))
        h = X.cpu().numpy()
        if kind == "shortA":
            LP = orc.project("T" if trans else "N", h, P)
        else:
            LP = orc.project("N" if trans else "T", h, P)
        del h
        _long_cache[key] = (X, P, LP)
    return _long_cache[key]


@pytest.mark.parametrize("prec,kind,mode,small", c3_cases(),
                         ids=[f"{p}-{k}-{m}-{s}" for p, k, m, s in c3_cases()])
def test_c3_skinny_full_size(orc, blas, dev, prec, kind, mode, small):
    import torch

    dt = torch.float64 if prec == "d" else torch.float32
    ta, tb = mode[0] == "T", mode[1] == "T"
    if kind == "shortA":
        M, N = small, LONG
        B, P, LP = long_operand(orc, blas, dev, prec, kind, tb)
        A = torch.empty((K3, M) if ta else (M, K3), dtype=dt, device=dev)
        blas.fill(A, M, K3, blas.NORMAL, 3, trans=ta)
    else:
        M, N = LONG, small
        A, P, LP = long_operand(orc, blas, dev, prec, kind, ta)
        B = torch.empty((N, K3) if tb else (K3, N), dtype=dt, device=dev)
        blas.fill(B, K3, N, blas.NORMAL, 3, trans=tb)
    C = torch.empty((M, N), dtype=dt, device=dev)
    blas.fill(C, M, N, blas.NORMAL, 4)
    # the band: columns of C for shortA, rows for thinB, placed by case
    idx = c3_cases().index((prec, kind, mode, small))
    start = [0, LONG // 2 - 40, LONG - BAND][idx % 3]
    C0_full = host(C)
    if kind == "shortA":
        C0 = host(C[:, start:start + BAND])
        opA = host(A.t() if ta else A)                                   # M x K
        opB = host(B[start:start + BAND, :].t() if tb else B[:, start:start + BAND])  # K x BAND
    else:
        C0 = host(C[start:start + BAND, :])
        opA = host(A[:, start:start + BAND].t() if ta else A[start:start + BAND, :])  # BAND x K
        opB = host(B.t() if tb else B)                                   # K x N
    blas.gemm(mode[0], mode[1], A, B, C, 1.0, 1.0, algo=blas.ALGO_SKINNY)
    torch.cuda.synchronize()
    got = host(C[:, start:start + BAND] if kind == "shortA" else C[start:start + BAND, :])
    ref = orc.gemm_xp("N", "N", opA, opB, C0, 1.0, 1.0)
    err, _ = orc.normwise(got, ref)
    tol = TOL64 if prec == "d" else tol32(K3)
    assert err <= tol, err
    # the whole C: R P = op(A) (op(B) P) + C0 P for shortA, R^T P = op(B)^T
    # (op(A)^T P) + C0^T P for thinB, with the long operand's share cached
    G = host(C)
    if kind == "shortA":
        RP = orc.project("T" if ta else "N", host(A), LP) + orc.project("N", C0_full, P)
        nw, line, i = orc.projected_errors(G, RP, P)
    else:
        RP = orc.project("N" if tb else "T", host(B), LP) + orc.project("T", C0_full, P)
        nw, line, i = orc.projected_errors(G, RP, P, side="left")
    assert nw <= tol and line <= (LINE64 if prec == "d" else 4 * tol), (nw, line, i)


# ---- c4 --------------------------------------------------------------------

N4, ROWS4 = 8192, 32
FORMS = {"fp32": tol32(N4), "bf16x6": 1e-6, "bf16x3": 3e-5, "fp16x3": 1e-6}


@pytest.fixture(scope="module")
def c4_operands(orc, blas, dev):
    """A, B, C of 8192^2 FP32 on the device for both distributions, their
    host copies, and the projection R X of R = A B + 0.5 C (X: 8192 x S,
    N(0,1), seed 44)."""
    import torch

    out = {}
    X = np.random.default_rng(44).standard_normal((N4, PROJ_S))
    for name, dist in (("uniform", blas.UNIFORM), ("loguniform", blas.LOGUNIF)):
        A = torch.empty((N4, N4), dtype=torch.float32, device=dev)
        B = torch.empty((N4, N4), dtype=torch.float32, device=dev)
        C = torch.empty((N4, N4), dtype=torch.float32, device=dev)
        blas.fill(A, N4, N4, dist, 3)
        blas.fill(B, N4, N4, dist, 4)
        blas.fill(C, N4, N4, blas.UNIFORM, 5)
        r0 = N4 // 2 + 8
        hA, hB, hC = host(A), host(B), host(C)
        RX = orc.gemm_projection("N", hA, "N", hB, hC, 1.0, 0.5, X)
        out[name] = (A, B, C, r0, np.ascontiguousarray(hA[r0:r0 + ROWS4]), hB,
                     np.ascontiguousarray(hC[r0:r0 + ROWS4]), X, RX)
        del hA, hC
    yield out
    out.clear()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dist", ["uniform", "loguniform"])
@pytest.mark.parametrize("form", ["fp32", "bf16x6", "bf16x3", "fp16x3", "fp8"])
def test_c4_precision_study(orc, blas, c4_operands, dist, form):
    import torch

    A, B, C, r0, hA, hB, hC, X, RX = c4_operands[dist]
    D = C.clone()
    alpha, beta = 1.0, 0.5
    if form == "fp32":
        blas.gemm("N", "N", A, B, D, alpha, beta)
    elif form.startswith("bf16"):
        blas.sgemm_split_bf16("N", "N", A, B, D, alpha, beta, nterms=int(form[-1]))
    elif form == "fp16x3":
        blas.sgemm_split_fp16("N", "N", A, B, D, alpha, beta)
    else:
        blas.sgemm_fp8("N", "N", A, B, D, alpha, beta)
    torch.cuda.synchronize()
    G = host(D)
    del D
    got = np.ascontiguousarray(G[r0:r0 + ROWS4])
    exact = orc.gemm_xp("N", "N", hA, hB, hC, alpha, beta)
    err, _ = orc.normwise(got, exact)
    nw, line, i = orc.projected_errors(G, RX, X)
    print(f"c4 {dist} {form}: normwise error vs FP64 {err:.3e} (band), {nw:.3e} (whole C, projected); "
          f"worst row {line:.3e} (row {i})")
    if form == "fp8":
        # the library's quantization of the band's rows and of all of B,
        # accumulated exactly: the band of the full quantized product
        q = orc.gemm_fp8("N", "N", hA, hB, hC, alpha, beta)
        qerr, _ = orc.normwise(got, q)
        assert qerr <= 2.0 ** -11, qerr
        assert 0.002 < err < 0.08, err
        assert 0.002 < nw < 0.08 and line < 4 * 0.08, (nw, line, i)
    else:
        assert err <= FORMS[form], err
        assert nw <= FORMS[form] and line <= 4 * FORMS[form], (nw, line, i)
