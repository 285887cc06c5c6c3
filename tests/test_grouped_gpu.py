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


def test_group_plan_two_streams_and_large_tiles(orc, blas, dev):
    """One plan run on two streams in turn (the second run is ordered after
    the first replay), on groups of FP32 and FP64 problems of every mode that
    take 128 tiles (M, N in 300..700), each bit for bit against the problem
    run alone; then a group of empty problems only."""
    import torch

    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    plan = blas.GroupPlan(16)
    try:
        groups = []
        for g, stream in enumerate((s1, s2, s1)):
            rng = np.random.default_rng(40 + g)
            probs = []
            for i in range(8):
                dt = np.float64 if i % 2 else np.float32
                mode = ["NN", "NT", "TN", "TT"][i % 4]
                M, N, K = (int(x) for x in rng.integers(300, 700, 3))
                A = orc.matrix(M, K, 0, 500 + 10 * g + i, dtype=dt, trans=mode[0] == "T")
                B = orc.matrix(K, N, 0, 600 + 10 * g + i, dtype=dt, trans=mode[1] == "T")
                C0 = orc.matrix(M, N, 0, 700 + 10 * g + i, dtype=dt)
                probs.append(dict(mode=mode, A=to_dev(A, dev), B=to_dev(B, dev), C0=to_dev(C0, dev)))
            outs = [p["C0"].clone() for p in probs]
            torch.cuda.current_stream().synchronize()
            descs = [blas.desc(p["mode"][0], p["mode"][1], p["A"], p["B"], c, 1.0, 0.5)
                     for p, c in zip(probs, outs)]
            plan.run(descs, stream=stream)
            groups.append((probs, outs))
        torch.cuda.synchronize()
        for g, (probs, outs) in enumerate(groups):
            for i, (p, c) in enumerate(zip(probs, outs)):
                alone = p["C0"].clone()
                blas.gemm(p["mode"][0], p["mode"][1], p["A"], p["B"], alone, 1.0, 0.5, algo=blas.ALGO_TILE,
                          splitk=1)
                torch.cuda.synchronize()
                assert torch.equal(c, alone), (g, i)
        empty = torch.zeros((0, 5), device=dev)
        plan.run([blas.desc("N", "N", empty, torch.zeros((5, 7), device=dev), torch.zeros((0, 7), device=dev))])
        torch.cuda.synchronize()
    finally:
        plan.close()


def test_group_plan_list_subsets(orc, blas, dev):
    """One plan replaying, back to back on one stream, groups that fill
    different sets of the four problem lists (FP32 only, FP64 only, 64 tiles
    only, 128 tiles only, one list each, all four), so each replay runs a
    different graph; every problem bit for bit against the problem alone."""
    import torch

    def problem(prec, big, seed):
        dt = np.float64 if prec == "d" else np.float32
        rng = np.random.default_rng(seed)
        mode = "NT"[rng.integers(2)] + "NT"[rng.integers(2)]
        M, N = (int(x) for x in rng.integers(384, 520, 2)) if big else (int(x) for x in rng.integers(5, 60, 2))
        K = int(rng.integers(1, 300))
        A = orc.matrix(M, K, 0, seed, dtype=dt, trans=mode[0] == "T")
        B = orc.matrix(K, N, 0, seed + 1, dtype=dt, trans=mode[1] == "T")
        C0 = orc.matrix(M, N, 0, seed + 2, dtype=dt)
        return dict(mode=mode, A=to_dev(A, dev), B=to_dev(B, dev), C0=to_dev(C0, dev))

    kinds = {"s64": ("s", False), "s128": ("s", True), "d64": ("d", False), "d128": ("d", True)}
    groups = [("s64", "s128"), ("d64", "d128"), ("s64", "d64"), ("s128", "d128"), ("d128",), ("s64",),
              ("s128", "d64"), ("s64", "s128", "d64", "d128")]
    plan = blas.GroupPlan(16)
    try:
        runs = []
        for g, kset in enumerate(groups):
            probs = [problem(*kinds[k], 1000 * g + 10 * i) for i, k in enumerate(kset * 2)]
            outs = [p["C0"].clone() for p in probs]
            plan.run([blas.desc(p["mode"][0], p["mode"][1], p["A"], p["B"], c, -0.5, 1.5)
                      for p, c in zip(probs, outs)])
            runs.append((probs, outs))
        torch.cuda.synchronize()
        for g, (probs, outs) in enumerate(runs):
            for i, (p, c) in enumerate(zip(probs, outs)):
                alone = p["C0"].clone()
                blas.gemm(p["mode"][0], p["mode"][1], p["A"], p["B"], alone, -0.5, 1.5, algo=blas.ALGO_TILE,
                          splitk=1)
                torch.cuda.synchronize()
                assert torch.equal(c, alone), (groups[g], i)
    finally:
        plan.close()


def test_group_wide_leading_dimension(orc, blas, dev):
    """A group holding an FP32 problem whose A rows are 4.4 million elements
    apart (past the grouped kernel's 32-bit load offsets: it runs through the
    per-call kernel on the plan's stream) beside an ordinary one; both against
    the oracle."""
    import torch

    M, K, N, lda = 3, 5000, 70, 4_400_000
    Abuf = torch.empty((M, lda), dtype=torch.float32, device=dev)
    A = Abuf[:, :K]
    B = torch.empty((K, N), dtype=torch.float32, device=dev)
    C = torch.empty((M, N), dtype=torch.float32, device=dev)
    blas.fill(A, M, K, 0, 61)
    blas.fill(B, K, N, 0, 62)
    blas.fill(C, M, N, 0, 63)
    A2 = torch.empty((200, 300), dtype=torch.float32, device=dev)
    C2 = torch.empty((200, N), dtype=torch.float32, device=dev)
    blas.fill(A2, 200, 300, 0, 64)
    blas.fill(C2, 200, N, 0, 65)
    c0, c20 = C.cpu().numpy().copy(), C2.cpu().numpy().copy()
    blas.gemm_grouped([blas.desc("N", "N", A, B, C, 1.0, 1.0), blas.desc("N", "N", A2, B[:300], C2, 1.0, 1.0)])
    torch.cuda.synchronize()
    a = np.ascontiguousarray(A.cpu().numpy())
    b = np.ascontiguousarray(B.cpu().numpy())
    ref = orc.gemm_xp("N", "N", a, b, c0, 1.0, 1.0)
    err, _ = orc.normwise(np.ascontiguousarray(C.cpu().numpy()), ref)
    assert err <= (np.sqrt(K) + 4) * 2.0 ** -23, err
    ref2 = orc.gemm_xp("N", "N", np.ascontiguousarray(A2.cpu().numpy()), b[:300], c20, 1.0, 1.0)
    err2, _ = orc.normwise(np.ascontiguousarray(C2.cpu().numpy()), ref2)
    assert err2 <= (np.sqrt(300) + 4) * 2.0 ** -23, err2
    del Abuf
    torch.cuda.empty_cache()


def test_group_rejects_bad_problem(blas, dev):
    import torch

    A = torch.zeros((4, 4), device=dev)
    d = blas.desc("N", "N", A, A, A)
    d.lda = 2  # lda < K
    with pytest.raises(blas.BlasdspError):
        blas.gemm_grouped([d])
