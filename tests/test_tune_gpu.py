"""Autotuner: the recorded choice is a real candidate, the AUTO dispatch and
the grouped launch follow it, and results stay within the oracle bar."""
import numpy as np
import pytest

from conftest import to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(16, 3000, 2048), (300, 20, 1500), (700, 600, 500)])
def test_autotune_then_dispatch(orc, blas, dev, shape):
    import torch

    M, N, K = shape
    blas.lib().blasdsp_tune_clear()
    try:
        for prec, dt in (("d", blas.FLOAT64), ("s", blas.FLOAT32)):
            c = blas.autotune(dt, "N", "T", M, N, K)
            assert c.algo in (blas.ALGO_TILE, blas.ALGO_SKINNY) and c.time_us > 0
            if c.algo == blas.ALGO_TILE:
                assert c.tile in (64, 128)
            else:
                assert min(M, N) <= 64 and 1 <= c.splitk <= 64
            q = blas.tune_query(dt, "N", "T", M, N, K)
            assert (q.algo, q.tile, q.splitk) == (c.algo, c.tile, c.splitk)
            npdt = np.float64 if prec == "d" else np.float32
            A = orc.matrix(M, K, 0, 1, dtype=npdt)
            B = orc.matrix(K, N, 0, 2, dtype=npdt, trans=True)
            C0 = orc.matrix(M, N, 0, 3, dtype=npdt)
            ref = orc.gemm_xp("N", "T", A, B, C0, 1.0, 0.5)
            dA, dB = to_dev(A, dev), to_dev(B, dev)
            for run in ("call", "grouped"):
                dC = to_dev(C0, dev)
                if run == "call":
                    blas.gemm("N", "T", dA, dB, dC, 1.0, 0.5)
                else:
                    blas.gemm_grouped([blas.desc("N", "T", dA, dB, dC, 1.0, 0.5)])
                torch.cuda.synchronize()
                err, _ = orc.normwise(np.ascontiguousarray(dC.cpu().numpy()), ref)
                assert err <= (1e-13 if prec == "d" else (np.sqrt(K) + 4) * 2.0 ** -23), (run, err)
    finally:
        blas.lib().blasdsp_tune_clear()
