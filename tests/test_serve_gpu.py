"""Continuous-batching request loop: results against the oracle, each result
identical alone and in the mix, memory-budget admission, rejection, stats."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def requests(blas, rng, n, max_dim=200, arrival_rate=None):
    out = []
    t = 0.0
    for i in range(n):
        dt = blas.FLOAT64 if rng.random() < 0.3 else blas.FLOAT32
        M, N, K = (int(x) for x in np.exp(rng.uniform(np.log(4), np.log(max_dim), 3)))
        if rng.random() < 0.3:
            M = int(rng.integers(8, 33))
        ta, tb = (b"NT"[rng.integers(2):][:1] for _ in range(2))
        if arrival_rate:
            t += rng.exponential(1.0 / arrival_rate)
        out.append(blas.Request(dt, ta, tb, M, N, K, float(rng.uniform(-1, 1)), float(rng.choice([0.0, 1.0])),
                                1000 + i, t))
    return out


def reference(orc, r):
    dt = np.float64 if r.dtype == 1 else np.float32
    ta, tb = r.transA.decode(), r.transB.decode()
    A = orc.matrix(r.K, r.M, 0, 3 * r.seed, dtype=dt) if ta == "T" else orc.matrix(r.M, r.K, 0, 3 * r.seed, dtype=dt)
    B = orc.matrix(r.N, r.K, 0, 3 * r.seed + 1, dtype=dt) if tb == "T" else orc.matrix(r.K, r.N, 0, 3 * r.seed + 1, dtype=dt)
    C = orc.matrix(r.M, r.N, 0, 3 * r.seed + 2, dtype=dt)
    return orc.gemm_xp(ta, tb, A, B, C, r.alpha, r.beta)


def tol(r):
    return 1e-13 if r.dtype == 1 else (np.sqrt(r.K) + 4) * 2.0 ** -23


def run(blas, reqs, grouped=1, budget=1 << 30, max_batch=64, inflight=2, scale=0.0, keep=True):
    outs = [np.empty((r.M, r.N), dtype=np.float64 if r.dtype == 1 else np.float32) for r in reqs] if keep else None
    cfg = blas.ServeConfig(budget, max_batch, inflight, grouped, scale)
    res, stats = blas.serve(reqs, cfg, outs)
    return res, stats, outs


def test_serve_results_and_alone_vs_mix(orc, blas, dev):
    rng = np.random.default_rng(4)
    reqs = requests(blas, rng, 40)
    res, stats, outs = run(blas, reqs)
    assert stats.completed == 40 and stats.rejected == 0
    assert stats.launches < 40  # batched
    for i, (r, x, o) in enumerate(zip(reqs, res, outs)):
        assert x.status == 0
        ref = reference(orc, r)
        err, _ = orc.normwise(o, ref)
        assert err <= tol(r), (i, err)
        assert x.checksum == pytest.approx(float(o.astype(np.float64).sum()), rel=1e-9, abs=1e-9)
    # each request alone (a batch of one) gives the same bits as in the mix
    for i in (0, 7, 19, 33):
        res1, _, outs1 = run(blas, [reqs[i]])
        np.testing.assert_array_equal(outs1[0], outs[i])
    # one launch per request: same results within the tolerance
    resp, statsp, outsp = run(blas, reqs, grouped=0)
    assert statsp.launches == 40
    for r, a, b in zip(reqs, outs, outsp):
        err, _ = orc.normwise(b, a.astype(np.float64))
        assert err <= 2 * tol(r)


def test_serve_memory_budget_and_rejection(orc, blas, dev):
    rng = np.random.default_rng(9)
    reqs = requests(blas, rng, 30, max_dim=400)
    sizes = [(8 if r.dtype == 1 else 4) * (r.M * r.K + r.K * r.N + r.M * r.N) for r in reqs]
    budget = int(max(sizes) * 1.5) + (1 << 16)
    big = blas.Request(1, b"N", b"N", 3000, 3000, 3000, 1.0, 0.0, 7, 0.0)  # > budget
    res, stats, _ = run(blas, reqs + [big], budget=budget, keep=False)
    assert res[-1].status == blas.ERR_NOMEM
    assert stats.rejected == 1 and stats.completed == 30
    assert stats.peak_resident_bytes <= budget
    assert all(x.status == 0 for x in res[:-1])


def test_serve_poisson_arrivals_stats(blas, dev):
    rng = np.random.default_rng(2)
    reqs = requests(blas, rng, 200, max_dim=128, arrival_rate=20000.0)
    res, stats, _ = run(blas, reqs, scale=1.0, keep=False)
    assert stats.completed == 200
    lat = sorted(x.latency_s for x in res)
    assert lat[0] >= 0
    assert stats.p50_s == pytest.approx(lat[99], rel=1e-12)
    assert stats.p99_s == pytest.approx(lat[197], rel=1e-12)
    assert all(x.done_s >= x.start_s for x in res)
    assert stats.gemms_per_s > 0 and stats.mean_batch >= 1
