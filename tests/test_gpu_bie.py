"""Sphere boundary-integral solve on the GPU (matrix-free FP64 GMRES, each
matvec one treecode evaluation) against a dense LU solve of the same discrete
system and against the Funk-Hecke modal solution of the continuous one."""
import numpy as np
import pytest

import bit_amd
from bit_amd import bie

pytestmark = pytest.mark.gpu


def test_matches_dense_solve(ctx):
    V, F = bie.icosphere(3)
    pts, area, h = bie.collocation(V, F)
    w = area / (4 * np.pi)
    kappa, d2 = 1.0, h * h
    self_coef = np.exp(-kappa * h) / h
    f = bie.random_rhs(6, seed=2)(pts)
    r = np.linalg.norm(pts[:, None, :] - pts[None, :, :], axis=2)
    M = bie.kernel(r, kappa, d2) * w[None, :]
    np.fill_diagonal(M, 0.5 + w * self_coef)
    ref = np.linalg.solve(M, f)
    s = bit_amd.System(ctx, pts[:, 0], pts[:, 1], pts[:, 2], w, 40)
    sig, it, rr = s.solve_second_kind(w, f, 0.5, self_coef, None, 30, 1e-11, 300, 12, 0.25,
                                      kappa, d2, None, "fp64")
    assert rr <= 1e-11
    assert np.linalg.norm(sig - ref) / np.linalg.norm(ref) < 1e-7


def test_converges_to_modal_solution(ctx):
    errs = []
    for level in (3, 4, 5):
        r = bie.solve_sphere(ctx, level=level, kappa=1.0, nterms=7, acc=0.6, restart=50,
                             tol=1e-8, lmax=6, seed=5, maxmemb=200)
        assert r["relres"] <= 1e-8
        assert r["iters"] < 60
        errs.append(r["err_vs_modal"])
    assert errs[0] > errs[1] > errs[2], errs
    assert errs[-1] < 2e-3, errs


def test_fp32_matvec_reaches_tolerance(ctx):
    r = bie.solve_sphere(ctx, level=4, kappa=1.0, nterms=7, acc=0.6, restart=50, tol=1e-6,
                         lmax=8, seed=6, maxmemb=200, precision="fp32")
    assert r["relres"] <= 1e-6


def test_resident_krylov_equals_host_gmres(ctx):
    """The device-resident solve (CGS2 Arnoldi in HBM) against the host GMRES
    core driven through the callback entry point with the same treecode
    operator; restart 4 forces several restarts."""
    V, F = bie.icosphere(3)
    pts, area, h = bie.collocation(V, F)
    w = area / (4 * np.pi)
    kappa, d2 = 1.0, h * h
    self_coef = float(np.exp(-kappa * h) / h)
    f = bie.random_rhs(6, seed=5)(pts)
    s = bit_amd.System(ctx, pts[:, 0], pts[:, 1], pts[:, 2], w, 40)
    args = (10, 0.3, kappa, d2, None, "fp64")

    def op(v):
        s.set_charges(w * v)
        return (0.5 + self_coef * w) * v + s.potential(*args)

    x0 = np.linspace(-1.0, 1.0, len(f))
    sig, it, rr = s.solve_second_kind(w, f, 0.5, self_coef, x0, 4, 1e-10, 400, *args)
    ref, it_h, rr_h = bit_amd.gmres(len(f), op, f, x0, restart=4, tol=1e-10, maxit=400)
    assert rr <= 1e-10 and rr_h <= 1e-10
    assert abs(it - it_h) <= 2, (it, it_h)
    assert np.linalg.norm(sig - ref) / np.linalg.norm(ref) < 1e-8
    # the true residual of the device solution
    res = np.linalg.norm(op(sig) - f) / np.linalg.norm(f)
    assert res < 2e-10
    # zero right-hand side and non-finite inputs
    z, it0, rr0 = s.solve_second_kind(w, np.zeros_like(f), 0.5, self_coef, None, 4, 1e-10, 10,
                                      *args)
    assert it0 == 0 and rr0 == 0.0 and not z.any()
    bad = f.copy()
    bad[3] = np.nan
    with pytest.raises(bit_amd.BitError):
        s.solve_second_kind(w, bad, 0.5, self_coef, None, 4, 1e-10, 10, *args)
