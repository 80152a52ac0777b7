"""Serving: mixed multi-system evaluations (one persistent launch per precision)
and the continuous-batching server give each request the same result it gets
alone."""
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


def test_persistent_launch_every_nterms(ctx):
    """One persistent launch per precision serves systems of every nterms (the
    far sweep chosen per item), sizes from one particle up, both kernels; the
    counters alternate between launches, so repeated launches must agree."""
    rng = np.random.default_rng(8)
    systems, params, alone = [], [], []
    sizes = [1, 5, 63, 64, 65, 700, 3000, 9000]
    for k, nt in enumerate(range(1, 14)):
        n = sizes[k % len(sizes)]
        P = workloads.make("blobs" if k % 2 else "uniform", n, 300 + k)
        kap = 0.0 if k % 3 == 0 else float(rng.uniform(0.2, 2.0))
        d2 = float(rng.uniform(0, 0.02)) ** 2
        for prec in ("fp32", "fp64"):
            s = bit_amd.System(ctx, P.x, P.y, P.z, P.q, int(rng.integers(8, 120)))
            alone.append(s.potential(nt, 0.6, kap, d2, None, prec))
            systems.append(s)
            params.append(bit_amd.make_params(nt, 0.6, kap, d2, None, prec))
    for rep in range(3):
        order = rng.permutation(len(systems))
        outs, _ = bit_amd.eval_multi(ctx, [systems[i] for i in order],
                                     [params[i] for i in order])
        for i, b in zip(order, outs):
            assert np.array_equal(alone[i], b), (rep, int(i))
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
    assert any(r.kind == "bie" for r in reqs)
    for r in reqs:
        got_iters = r.iters
        ref = serving.solve_alone(ctx, r)
        assert np.array_equal(r.result, ref), (r.rid, r.kind)
        if r.kind == "bie":
            assert r.relres <= r.tol and r.iters == got_iters


def test_budget_admission_serialises_but_completes():
    reqs = serving.make_requests(12, rate=5000.0, seed=10, nmin=20000, nmax=40000,
                                 bie_fraction=0.0)
    budget = max(r.est_bytes for r in reqs) + 1  # at most one request in flight
    srv = serving.Server(builders=2, hbm_budget=budget)
    m = srv.run(reqs)
    srv.close()
    assert not any("error" in r.extra for r in reqs)
    assert m["mean_systems_per_round"] == 1.0


def test_solves_stepped_together_equal_alone(ctx):
    """Several resident-Krylov solves of different sizes and parameters, stepped
    together in serving rounds that also carry plain evaluations, give the same
    iterates, iteration counts and residuals as each solve alone."""
    from bit_amd import bie
    cases = [(3, 1.0, "fp64", 7, 0.6), (4, 0.5, "fp32", 5, 0.5), (3, 2.0, "fp32", 9, 0.4),
             (5, 1.0, "fp32", 7, 0.6)]
    probs = []
    for lv, kap, prec, nt, acc in cases:
        V, F = bie.icosphere(lv)
        pts, area, h = bie.collocation(V, F)
        w = area / (4 * np.pi)
        f = bie.random_rhs(6, seed=lv + int(10 * kap))(pts)
        tol = 1e-9 if prec == "fp64" else 1e-6
        probs.append((pts, w, f, float(np.exp(-kap * h) / h), tol,
                      (nt, acc, kap, h * h, None, prec)))
    alone = []
    for pts, w, f, sc, tol, args in probs:
        s = bit_amd.System(ctx, pts[:, 0], pts[:, 1], pts[:, 2], w, 64)
        alone.append(s.solve_second_kind(w, f, 0.5, sc, None, 8, tol, 200, *args))
        s.close()
    systems, solves = [], []
    for pts, w, f, sc, tol, args in probs:
        s = bit_amd.System(ctx, pts[:, 0], pts[:, 1], pts[:, 2], w, 64)
        systems.append(s)
        solves.append(bit_amd.Krylov(s, w, f, 0.5, sc, None, 8, tol, 200, *args))
    P = workloads.gaussian_blobs(20000, seed=77)
    ev = bit_amd.System(ctx, P.x, P.y, P.z, P.q, 100)
    eargs = (7, 0.5, 0.8, 1e-4, None, "fp32")
    want_eval = ev.potential(*eargs)
    rounds = 0
    while not all(k.done for k in solves):
        outs, ms = bit_amd.serve_round(ctx, [ev], [bit_amd.make_params(*eargs)], solves)
        assert np.array_equal(outs[0], want_eval)
        rounds += 1
        assert rounds < 500
    for k, (sig, it, rr) in zip(solves, alone):
        done, it2, rr2 = k.status()
        assert done == 1 and it2 == it and rr2 == rr
        assert np.array_equal(k.solution(), sig)
        k.close()
    for s in systems + [ev]:
        s.close()
