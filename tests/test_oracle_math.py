"""The oracle against what the mathematics fixes: an independent power-series
expansion of the kernel, closed-form direct sums in extended precision,
brute-force moments, and convergence of the treecode to the direct sum."""
import numpy as np
import pytest

from bit_amd import workloads
from conftest import rel_l2, rel_max
from series import taylor_coeffs


@pytest.mark.parametrize("kappa,delta2", [(0.0, 0.0), (0.7, 0.0), (2.0, 0.03), (0.0, 0.01)])
@pytest.mark.parametrize("x", [(0.9, -0.4, 0.3), (-0.2, 1.3, 0.05), (0.5, 0.5, -0.8)])
def test_taylor_coeffs_vs_power_series(oracle, kappa, delta2, x):
    P = 8
    a = oracle.taylor_coeffs(P + 1, x[0], x[1], x[2], x[0] ** 2 + x[1] ** 2 + x[2] ** 2 + delta2,
                             kappa)
    ref = taylor_coeffs(P, x, kappa, delta2)
    i, j, k = np.indices((P + 1,) * 3)
    m = (i + j + k) <= P
    err = np.abs(a[:P + 1, :P + 1, :P + 1][m] - ref[m])
    scale = np.abs(ref[m]).max()
    assert err.max() <= 1e-12 * scale


def test_taylor_series_converges_to_kernel(oracle):
    """sum_k a_k y^k -> G(x - y) for a source offset y inside the radius."""
    x = np.array([0.8, -0.5, 0.4])
    y = np.array([0.05, 0.08, -0.06])
    kappa, d2 = 1.1, 0.002
    R = np.sqrt(np.sum((x - y) ** 2) + d2)
    G = np.exp(-kappa * R) / R
    errs = []
    for nt in (3, 6, 9, 12):
        a = oracle.taylor_coeffs(nt, *x, float(np.sum(x ** 2) + d2), kappa)
        i, j, k = np.indices(a.shape)
        m = (i + j + k) < nt
        errs.append(abs(np.sum(a[m] * (y[0] ** i * y[1] ** j * y[2] ** k)[m]) - G) / G)
    assert all(e2 < e1 for e1, e2 in zip(errs, errs[1:]))
    assert errs[-1] < 1e-10


def test_direct_sum_closed_form(oracle):
    P = workloads.uniform_cube(60, seed=7)
    kappa, d2 = 0.8, 0.0004
    got = oracle.direct(P.x, P.y, P.z, P.q, kappa, d2)
    pos = np.stack([P.x, P.y, P.z], 1).astype(np.longdouble)
    ref = np.zeros(P.n, dtype=np.longdouble)
    for i in range(P.n):
        r = np.sqrt(np.sum((pos - pos[i]) ** 2, axis=1) + np.longdouble(d2))
        v = P.q.astype(np.longdouble) * np.exp(-np.longdouble(kappa) * r) / r
        v[i] = 0  # self excluded
        ref[i] = v.sum()
    assert rel_max(got, ref.astype(np.float64)) < 1e-13


def test_moments_brute_force(oracle):
    P = workloads.gaussian_blobs(3000, seed=8)
    t = oracle.tree(P.x, P.y, P.z, 50)
    e = t.export()
    nt = 6
    m = t.moments(P.q, nt)
    for node in (0, 3, e.numtree - 1):
        mem = e.node_members(node)
        b = e.boxes[node]
        dx = (P.x[mem] - b[6]).astype(np.longdouble)
        dy = (P.y[mem] - b[7]).astype(np.longdouble)
        dz = (P.z[mem] - b[8]).astype(np.longdouble)
        q = P.q[mem].astype(np.longdouble)
        for i in range(nt):
            for j in range(nt - i):
                for k in range(nt - i - j):
                    ref = np.sum(q * dx ** i * dy ** j * dz ** k)
                    tol = 1e-13 * float(np.sum(np.abs(q) * (np.abs(dx) + np.abs(dy) + np.abs(dz) + 1e-300) ** (i + j + k)))
                    assert abs(m[node, i, j, k] - float(ref)) <= tol


@pytest.mark.parametrize("kappa,delta2", [(0.0, 0.0), (0.5, 0.0001)])
def test_treecode_converges_to_direct_sum(oracle, kappa, delta2):
    P = workloads.gaussian_blobs(6000, seed=9)
    t = oracle.tree(P.x, P.y, P.z, 100)
    tidx = workloads.sample_targets(P.n, 300, seed=1)
    ds = oracle.direct(P.x, P.y, P.z, P.q, kappa, delta2, tidx)
    errs = []
    for nt in (2, 4, 6, 8, 10, 12):
        m = t.moments(P.q, nt)
        phi = t.potential(P.q, m, nt, 0.25, kappa, delta2, delta2, tidx)
        errs.append(rel_l2(phi, ds))
    assert all(e2 < e1 for e1, e2 in zip(errs, errs[1:]))
    assert errs[-1] < 1e-6


def test_field_is_minus_gradient_of_potential(oracle):
    """E from the field traversal equals -grad phi of the regularized Coulomb
    direct sum (central differences), and the tree field converges to it."""
    P = workloads.uniform_cube(2000, seed=10)
    d2 = 0.01
    tidx = np.arange(0, 2000, 97, dtype=np.int32)
    Ed = oracle.direct_force(P.x, P.y, P.z, P.q, d2, tidx)
    h = 1e-5
    pos = np.stack([P.x, P.y, P.z], 1)
    for r, i in enumerate(tidx[:5]):
        g = []
        for e in range(3):
            vals = []
            for sgn in (1, -1):
                p = pos[i].copy()
                p[e] += sgn * h
                d = pos - p
                R = np.sqrt(np.sum(d ** 2, 1) + d2)
                v = P.q / R
                v[i] = 0
                vals.append(v.sum())
            g.append(-(vals[0] - vals[1]) / (2 * h))
        assert np.allclose(Ed[r], g, rtol=1e-6, atol=1e-8 * np.abs(g).max())
    t = oracle.tree(P.x, P.y, P.z, 60)
    errs = []
    for nt in (4, 8, 12):
        m = t.moments(P.q, nt)
        errs.append(rel_l2(t.treeforce(P.q, m, nt, 0.25, d2, d2, tidx), Ed))
    assert errs[0] > errs[1] > errs[2]
    assert errs[-1] < 3e-5


def test_point_targets(oracle):
    """Point evaluation (p = -1): at a particle position it equals the
    particle evaluation plus the target's own term q_i G(0) (kappa 0,
    delta^2 > 0); at arbitrary points it converges to a numpy direct sum; a
    point on a particle with delta 0 skips that particle."""
    P = workloads.gaussian_blobs(3000, seed=9)
    t = oracle.tree(P.x, P.y, P.z, 60)
    d2 = 0.01
    m = t.moments(P.q, 8)
    tidx = np.arange(0, P.n, 7)
    phi = t.potential(P.q, m, 8, 0.5, 0.0, d2, d2, tidx)
    pts = t.potential_points(P.q, m, 8, 0.5, 0.0, d2, d2, P.x[tidx], P.y[tidx], P.z[tidx])
    assert rel_max(pts, phi + P.q[tidx] / np.sqrt(d2)) < 1e-13
    rng = np.random.default_rng(4)
    X = rng.uniform(-1.5, 1.5, size=(200, 3))
    X[:5] = np.stack([P.x[:5], P.y[:5], P.z[:5]], axis=1)
    for kappa, dd in ((0.7, 0.0), (0.0, 1e-4)):
        r = np.sqrt(((X[:, None, :] - np.stack([P.x, P.y, P.z], 1)[None]) ** 2).sum(-1) ** 1 + dd)
        G = np.where(r * r > 1e-14, np.exp(-kappa * r) / np.where(r > 0, r, 1.0), 0.0)
        ref = G @ P.q
        mm = t.moments(P.q, 12)
        got = t.potential_points(P.q, mm, 12, 0.3, kappa, dd, max(dd, 0.01), *X.T.copy())
        assert rel_l2(got, ref) < 1e-6
    Ed = np.zeros((200, 3))
    S = np.stack([P.x, P.y, P.z], 1)
    for i in range(200):
        d = X[i] - S
        r2 = (d * d).sum(1) + 0.01
        Ed[i] = (P.q / r2 ** 1.5) @ d
    errs = []
    for nt in (4, 8, 12):
        mm = t.moments(P.q, nt)
        errs.append(rel_l2(t.treeforce_points(P.q, mm, nt, 0.25, 0.01, 0.01, *X.T.copy()), Ed))
    assert errs[0] > errs[1] > errs[2] and errs[2] < 3e-5
    # at a particle position the point form adds the particle's own term,
    # whose regularized gradient is exactly zero
    mm = t.moments(P.q, 6)
    Ep = t.treeforce_points(P.q, mm, 6, 0.5, 0.01, 0.01, P.x[tidx], P.y[tidx], P.z[tidx])
    assert rel_max(Ep, t.treeforce(P.q, mm, 6, 0.5, 0.01, 0.01, tidx)) < 1e-13


@pytest.mark.parametrize("l", [0, 1, 2, 3])
@pytest.mark.parametrize("x,d2", [((0.9, -0.4, 0.3), 0.01), ((-0.2, 1.3, 0.05), 0.2),
                                  ((0.5, 0.5, -0.8), 0.0)])
def test_powinv_coeffs_vs_power_series(oracle, l, x, d2):
    """Coefficients of R^-(2l+1) (terms of the regularized field family)
    against the binomial series of (1 + w)^-(2l+1)/2; l = 0 also against the
    kappa = 0 recurrence used for the potential."""
    from series import powinv_coeffs
    P = 8
    R2 = x[0] ** 2 + x[1] ** 2 + x[2] ** 2 + d2
    a = oracle.powinv_coeffs(P + 1, l, *x, R2)
    ref = powinv_coeffs(P, x, l, d2)
    i, j, k = np.indices((P + 1,) * 3)
    m = (i + j + k) <= P
    assert np.abs(a[:P + 1, :P + 1, :P + 1][m] - ref[m]).max() <= 1e-12 * np.abs(ref[m]).max()
    if l == 0:
        t = oracle.taylor_coeffs(P + 1, *x, R2, 0.0)
        assert np.abs(a - t).max() <= 1e-13 * np.abs(t).max()


def _family_phi(P, X, d2, L):
    """Potential of the family G_L = sum_l w_l d2^l R^-(2l+1) at points X."""
    S = np.stack([P.x, P.y, P.z], 1)
    out = np.zeros(len(X))
    for i, xi in enumerate(X):
        R2 = ((xi - S) ** 2).sum(1) + d2
        g = np.zeros_like(R2)
        w = 1.0
        for l in range(L + 1):
            if l:
                w *= (l - 0.5) / l
            g += w * d2 ** l * R2 ** (-(2 * l + 1) / 2.0)
        out[i] = g @ P.q
    return out


@pytest.mark.parametrize("L", [1, 2, 3])
def test_field_family(oracle, L):
    """KERNEL_ORDER >= 1: the direct field is -grad of the family's potential
    (central differences, the field taken at points off the particles), it
    grows with L as every added term is positive, and the treecode field
    converges to it as nterms grows."""
    P = workloads.gaussian_blobs(1500, seed=13)
    d2 = 0.01
    rng = np.random.default_rng(L)
    X = rng.uniform(-0.6, 0.6, size=(6, 3))
    # a point evaluation via a one-particle "target" appended with charge 0
    xs = np.concatenate([P.x, X[:, 0]])
    ys = np.concatenate([P.y, X[:, 1]])
    zs = np.concatenate([P.z, X[:, 2]])
    qs = np.concatenate([P.q, np.zeros(len(X))])
    tidx = np.arange(P.n, P.n + len(X))
    E = oracle.direct_force(xs, ys, zs, qs, d2, tidx, kernel_order=L)
    h = 1e-5
    for e in range(3):
        dX = np.zeros_like(X)
        dX[:, e] = h
        g = -(_family_phi(P, X + dX, d2, L) - _family_phi(P, X - dX, d2, L)) / (2 * h)
        assert np.allclose(E[:, e], g, rtol=1e-6, atol=1e-7 * np.abs(g).max())
    tid = workloads.sample_targets(P.n, 300, seed=2)
    E0 = oracle.direct_force(P.x, P.y, P.z, P.q, d2, tid, kernel_order=0)
    EL = oracle.direct_force(P.x, P.y, P.z, P.q, d2, tid, kernel_order=L)
    assert rel_l2(EL, E0) > 1e-4
    t = oracle.tree(P.x, P.y, P.z, 40)
    errs = []
    for nt in (3, 6, 9):
        m = t.moments(P.q, nt)
        errs.append(rel_l2(t.treeforce(P.q, m, nt, 0.3, d2, d2, tid, kernel_order=L), EL))
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-4
