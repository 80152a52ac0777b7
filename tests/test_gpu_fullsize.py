"""Parity at the benchmark configurations' full sizes.

* 4M-particle 16-blob screened workload at its tuned point (kappa 0.5, delta
  0.005, p 11, acc 0.255, N0 500): tree (device build) and deep moments equal the
  oracle's, the FP64 potential is within 1e-10 of the oracle's traversal on
  2,000 sampled targets, FP32 within FP32_TOL of the same.
* 1M uniform, regularized Coulomb (delta 0.01, p 8, acc 0.5, N0 1000), FP64
  and FP32, the same bars; the error of both against the FP64 direct sum on
  10,000 targets is printed.
* Sphere BIE at icosphere level 7 (327,680 unknowns; kappa 1, delta = mean
  edge, acc 0.6, p 6, GMRES(50), tol 1e-6): converged, and closer to the modal
  solution of the continuous equation than level 5.
"""
import numpy as np
import pytest

import bit_amd
from bit_amd import bie, workloads
from conftest import rel_l2, rel_max

pytestmark = pytest.mark.gpu
FP32_TOL = 2e-6


def _check_tree(s, t_ora):
    t = s.export_tree()
    o = t_ora.export()
    assert t.numtree == o.numtree
    assert np.array_equal(t.children, o.children)
    assert np.array_equal(t.boxes.view(np.uint64), o.boxes.view(np.uint64))
    size = (t.end - t.begin).astype(np.int64)
    assert np.array_equal(size, o.nummemb.astype(np.int64))


def _full(ctx, oracle, P, maxmemb, nterms, acc, kappa, d2, ntargets, seed, tol32=FP32_TOL):
    dev = [bit_amd.DeviceArray.from_numpy(ctx, a) for a in (P.x, P.y, P.z, P.q)]
    s = bit_amd.System(ctx, *dev, maxmemb)
    try:
        p64 = s.potential(nterms, acc, kappa, d2, None, "fp64")
        p32 = s.potential(nterms, acc, kappa, d2, None, "fp32")
        s.moments(nterms)
        mom = s.moments_cube()
        t_ora = oracle.tree(P.x, P.y, P.z, maxmemb)
        _check_tree(s, t_ora)
        m_ora = t_ora.moments(P.q, nterms)
        # every node's moments (the GPU's upward pass, the deepest heights
        # included) against the oracle's sums over the node's particles,
        # relative to the node's largest moment
        nt = mom.shape[0]
        scale = np.abs(m_ora).reshape(nt, -1).max(axis=1)
        diff = np.abs(mom - m_ora).reshape(nt, -1).max(axis=1)
        worst = float(np.max(diff / np.maximum(scale, 1e-300)))
        print(f"n={P.n}: {nt} nodes, moments worst node rel-max {worst:.2e}")
        assert worst < 1e-11
        tidx = workloads.sample_targets(P.n, ntargets, seed=seed)
        ref = t_ora.potential(P.q, m_ora, nterms, acc, kappa, d2, d2, tidx)
        e64 = rel_max(p64[tidx], ref)
        e32 = rel_l2(p32[tidx], ref)
        print(f"n={P.n}: fp64 rel-max vs oracle {e64:.2e}, fp32 rel-L2 {e32:.2e}")
        assert e64 < 1e-10
        assert e32 < tol32
        return s, dev, p64, p32
    except Exception:
        s.close()
        for a in dev:
            a.free()
        raise


def test_blobs4m_fp64_and_fp32_equal_oracle(ctx, oracle):
    P = workloads.gaussian_blobs(4_000_000, seed=3)
    s, dev, p64, p32 = _full(ctx, oracle, P, 500, 12, 0.255, 0.5, 0.005 ** 2, 2000, seed=17)
    s.close()
    for a in dev:
        a.free()


def test_uniform1m_regularized_fp64_and_fp32(ctx, oracle):
    P = workloads.uniform_cube(1_000_000, seed=2)
    d2 = 0.01 ** 2
    s, dev, p64, p32 = _full(ctx, oracle, P, 1000, 9, 0.5, 0.0, d2, 2000, seed=18)
    try:
        tidx = workloads.sample_targets(P.n, 10_000, seed=7)
        ref = s.direct(0.0, d2, tidx)
        e64, e32 = rel_l2(p64[tidx], ref), rel_l2(p32[tidx], ref)
        print(f"uniform 1M, delta 0.01, p 8, acc 0.5, N0 1000: rel-L2 vs FP64 direct sum on "
              f"10,000 targets: fp64 {e64:.3e}, fp32 {e32:.3e}")
        assert e64 < 1e-3 and abs(e32 - e64) < 1e-5  # truncation error, same in both modes
    finally:
        s.close()
        for a in dev:
            a.free()


@pytest.fixture(scope="module")
def bie_tight():
    """Levels 5 and 7 with a tight treecode (p 10, acc 0.3, tol 1e-9): their
    distance to the modal solution is the discretisation error alone."""
    import bit_amd as b
    c = b.Context(0)
    out = {lv: bie.solve_sphere(c, level=lv, kappa=1.0, nterms=11, acc=0.3, restart=50,
                                tol=1e-9, lmax=8, seed=17, maxmemb=500, precision="fp64")
           for lv in (5, 7)}
    c.close()
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sphere_bie_level7(ctx, precision, bie_tight):
    r7 = bie.solve_sphere(ctx, level=7, kappa=1.0, nterms=7, acc=0.6, restart=50, tol=1e-6,
                          lmax=8, seed=17, maxmemb=500, precision=precision)
    t5, t7 = bie_tight[5], bie_tight[7]
    dist = np.linalg.norm(r7["sigma"] - t7["sigma"]) / np.linalg.norm(t7["sigma"])
    print(f"{precision}: level 7 n={r7['n']} iters {r7['iters']} relres {r7['relres']:.2e} "
          f"err vs modal {r7['err_vs_modal']:.2e}, vs tight solve {dist:.2e}; tight: level 5 "
          f"{t5['err_vs_modal']:.2e}, level 7 {t7['err_vs_modal']:.2e}; solve "
          f"{r7['solve_s']:.3f}s")
    assert r7["n"] == 327_680
    assert r7["relres"] <= 1e-6 and r7["iters"] < 50
    # the discretisation converges towards the continuous solution ...
    assert t7["relres"] <= 1e-9 and t7["err_vs_modal"] < 0.5 * t5["err_vs_modal"]
    # ... and at the benchmark's treecode point (acc 0.6, p 6) the solve stays
    # within the treecode's truncation error of the tight solve
    assert dist < 5e-4
    assert r7["err_vs_modal"] < t7["err_vs_modal"] + dist + 1e-6
