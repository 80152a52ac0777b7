"""Serving: mixed multi-system launches and the continuous-batching server give
each request the same result it gets alone."""
import numpy as np
import pytest

import bit_amd
from bit_amd import serving, workloads

pytestmark = pytest.mark.gpu


def test_mixed_launch_bit_identical(ctx):
    rng = np.random.default_rng(4)
    systems, params, alone = [], [], []
    for k in range(12):
        n = int(rng.integers(30, 30000))
        P = workloads.make("blobs" if k % 2 else "uniform", n, 100 + k)
        nt = int(rng.choice([5, 7, 9]))
        kap = float(rng.choice([0.0, 0.7, 1.9]))
        d2 = float(rng.uniform(0, 0.02)) ** 2
        prec = "fp32" if k % 3 else "fp64"
        s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, int(rng.integers(8, 200)))
        alone.append(s.potential(nt, 0.5, kap, d2, None, prec))
        systems.append(s)
        params.append(bit_amd.make_params(nt, 0.5, kap, d2, None, prec))
    outs, ms = bit_amd.eval_multi(ctx, systems, params)
    for a, b in zip(alone, outs):
        assert np.array_equal(a, b)
    # a different mix and order
    idx = [5, 0, 11, 3]
    outs2, _ = bit_amd.eval_multi(ctx, [systems[i] for i in idx], [params[i] for i in idx])
    for i, b in zip(idx, outs2):
        assert np.array_equal(alone[i], b)
    for s in systems:
        s.close()


def test_server_results_equal_alone(ctx):
    reqs = serving.make_requests(40, rate=2000.0, seed=9, nmin=500, nmax=20000)
    for r in reqs:
        if r.kind == "bie":
            r.level = min(r.level, 4)
            r.n = 20 * 4 ** r.level
    srv = serving.Server(builders=4, hbm_budget=1 << 30)
    m = srv.run(reqs)
    srv.close()
    assert not any("error" in r.extra for r in reqs)
    assert m["requests"] == 40 and m["p99_ms"] >= m["p50_ms"] > 0
    for r in reqs:
        ref = serving.solve_alone(ctx, r)
        if r.kind == "eval":
            assert np.array_equal(r.result, ref), r.rid
        else:
            assert r.relres <= r.tol
            assert np.max(np.abs(r.result - ref)) <= 1e-12 * np.max(np.abs(ref)), r.rid


def test_budget_admission_serialises_but_completes():
    reqs = serving.make_requests(12, rate=5000.0, seed=10, nmin=20000, nmax=40000,
                                 bie_fraction=0.0)
    budget = max(r.est_bytes for r in reqs) + 1  # at most one request in flight
    srv = serving.Server(builders=2, hbm_budget=budget)
    m = srv.run(reqs)
    srv.close()
    assert not any("error" in r.extra for r in reqs)
    assert m["mean_systems_per_round"] == 1.0
