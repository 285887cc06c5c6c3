"""Grouped persistent GEMM: every problem of a mixed group (precisions, modes,
shapes from 1 to a few hundred, skinny, empty, alpha = 0) against the oracle
and bit for bit against the same problem run alone through the tile kernel;
the hipGraph plan replayed on different groups."""
import numpy as np
import pytest

from conftest import to_dev

pytestmark = pytest.mark.gpu


def make_group(orc, dev, rng, n):
    """Random problems with padded leading dimensions; returns (problems, refs)."""
    probs = []
    for i in range(n):
        prec = "d" if rng.random() < 0.3 else "s"
        dt = np.float64 if prec == "d" else np.float32
        mode = "NT"[rng.integers(2)] + "NT"[rng.integers(2)]
        M, N, K = (int(x) for x in np.exp(rng.uniform(0, np.log(300), 3)))
        if rng.random() < 0.2:
            M = int(rng.integers(1, 33))
        if rng.random() < 0.1:
            K = 0
        alpha = 0.0 if rng.random() < 0.05 else float(rng.uniform(-2, 2))
        beta = float(rng.choice([0.0, 1.0, -0.5]))
        A = orc.matrix(M, K, 0, 100 + i, dtype=dt, ld=(M if mode[0] == "T" else max(K, 1)) + 1,
                       trans=mode[0] == "T", pad=0.0)
        B = orc.matrix(K, N, 0, 200 + i, dtype=dt, ld=(K if mode[1] == "T" else N) + 2,
                       trans=mode[1] == "T", pad=0.0)
        C0 = orc.matrix(M, N, 0, 300 + i, dtype=dt, ld=N + 3, pad=0.0)
        ref = orc.gemm_xp(mode[0], mode[1], A, B, C0, alpha, beta)
        probs.append(dict(prec=prec, mode=mode, A=to_dev(A, dev), B=to_dev(B, dev), C0=to_dev(C0, dev),
                          alpha=alpha, beta=beta, ref=ref, K=K))
    return probs


@pytest.mark.parametrize("seed", [0, 1])
def test_grouped_matches_alone_and_oracle(orc, blas, dev, seed):
    import torch

    rng = np.random.default_rng(seed)
    probs = make_group(orc, dev, rng, 60)
    outs = [p["C0"].clone() for p in probs]
    alone = [p["C0"].clone() for p in probs]
    descs = [blas.desc(p["mode"][0], p["mode"][1], p["A"], p["B"], c, p["alpha"], p["beta"])
             for p, c in zip(probs, outs)]
    blas.gemm_grouped(descs)
    for p, c in zip(probs, alone):
        blas.gemm(p["mode"][0], p["mode"][1], p["A"], p["B"], c, p["alpha"], p["beta"], algo=blas.ALGO_TILE,
                  splitk=1)
    torch.cuda.synchronize()
    for i, (p, c, a) in enumerate(zip(probs, outs, alone)):
        got = np.ascontiguousarray(c.cpu().numpy())
        assert np.array_equal(got, a.cpu().numpy()), f"problem {i}: grouped differs from alone"
        err, _ = orc.normwise(got, p["ref"])
        tol = 1e-13 if p["prec"] == "d" else (np.sqrt(max(p["K"], 1)) + 4) * 2.0 ** -23
        assert err <= tol, (i, err)


def test_group_plan_replays(orc, blas, dev):
    import torch

    plan = blas.GroupPlan(128)
    try:
        for seed in (5, 6, 7):
            rng = np.random.default_rng(seed)
            probs = make_group(orc, dev, rng, 25 + seed)
            outs = [p["C0"].clone() for p in probs]
            descs = [blas.desc(p["mode"][0], p["mode"][1], p["A"], p["B"], c, p["alpha"], p["beta"])
                     for p, c in zip(probs, outs)]
            plan.run(descs)
            plan.run(descs[:3])  # a second, smaller group right behind: C of the first three updated twice
            torch.cuda.synchronize()
            for i, (p, c) in enumerate(zip(probs, outs)):
                got = np.ascontiguousarray(c.cpu().numpy())
                if i < 3:
                    continue
                err, _ = orc.normwise(got, p["ref"])
                tol = 1e-13 if p["prec"] == "d" else (np.sqrt(max(p["K"], 1)) + 4) * 2.0 ** -23
                assert err <= tol, (seed, i, err)
    finally:
        plan.close()


def test_group_rejects_bad_problem(blas, dev):
    import torch

    A = torch.zeros((4, 4), device=dev)
    d = blas.desc("N", "N", A, A, A)
    d.lda = 2  # lda < K
    with pytest.raises(blas.BlasdspError):
        blas.gemm_grouped([d])
