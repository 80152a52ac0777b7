"""The FP32 far kernel's radial-derivative form against the oracle's Taylor sum.

The FP32 far kernel (k_farq) evaluates sum_{|k|<=p} M_k a_k as
    (1/R) sum_t z^t sum_{|l|<=p-2t} ghat_{|l|+t} Q_{|l|+t,l} w^l
with w = rho u, z = rho^2, ghat_m = R^(2m+1) (R^-1 d/dR)^m G and the
per-cluster Q of k_qcoef.  This restates that form in FP64, in the kernel's
sweep order (ratio Horner), and checks it against the oracle's coefficient
recurrence (ora_taylor_coeffs, pinned to the reference) on random clusters
and targets that pass the MAC, screened and Coulomb, with and without delta.
It also checks the consumption-order count nqcoef against bench.py's copy.
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import Oracle


def monos(p):
    return [(n - m, m - k, k) for n in range(p + 1) for m in range(n + 1) for k in range(m + 1)]


def qcoef(mhat, p):
    """Q_{m,l} = (-1)^|l| / l! * sum_{|j|=m-|l|} Mhat_{l+2j} / (j! 2^|j|)."""
    f = math.factorial
    q = {}
    for l in monos(p):
        d = sum(l)
        for t in range((p - d) // 2 + 1):
            s = 0.0
            for j0 in range(t + 1):
                for j1 in range(t - j0 + 1):
                    j2 = t - j0 - j1
                    k = (l[0] + 2 * j0, l[1] + 2 * j1, l[2] + 2 * j2)
                    s += mhat[k] / (f(j0) * f(j1) * f(j2) * 2 ** t)
            q[(d + t, l)] = (-1) ** d * s / (f(l[0]) * f(l[1]) * f(l[2]))
    return q


def far_q(q, p, d, s, kappa, delta2):
    """(1/R) sum_t z^t ghat_t Phi_t, Phi_t by the ratio Horner sweep of k_farq:
    t descending, then l0, l1, l2 descending, a step in direction e at degree
    n multiplying by r_{n+t} w_e, r_m = ghat_{m+1} / ghat_m."""
    R = math.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2 + delta2)
    lam, rho = kappa * R, s / R
    w = [rho * d[e] / R for e in range(3)]
    z = rho * rho
    r = [-(1.0 + lam)]
    for m in range(1, p):
        r.append(lam * lam / r[m - 1] - (2 * m + 1))
    g = [math.exp(-lam)]
    for t in range(p // 2):
        g.append(g[t] * r[t])
    S = 0.0
    for t in range(p // 2, -1, -1):
        L = p - 2 * t
        phi = 0.0
        for l0 in range(L, -1, -1):
            a = 0.0
            for l1 in range(L - l0, -1, -1):
                b = 0.0
                for l2 in range(L - l0 - l1, -1, -1):
                    qv = q[(l0 + l1 + l2 + t, (l0, l1, l2))]
                    b = qv if l2 == L - l0 - l1 else b * r[l0 + l1 + l2 + t] * w[2] + qv
                a = b if l1 == L - l0 else a * r[l0 + l1 + t] * w[1] + b
            phi = a if l0 == L else phi * r[l0 + t] * w[0] + a
        S = g[t] * phi if t == p // 2 else S * z + g[t] * phi
    return S / R


@pytest.fixture(scope="module")
def ora():
    return Oracle()


@pytest.mark.parametrize("p,kappa,delta2", [(0, 0.5, 0.0), (1, 0.5, 0.0), (5, 0.0, 0.0),
                                            (8, 0.5, 2.5e-5), (11, 0.5, 2.5e-5),
                                            (11, 2.0, 4e-4), (12, 0.0, 1e-4)])
def test_radial_form_equals_taylor_sum(ora, p, kappa, delta2):
    rng = np.random.default_rng(100 + p)
    worst = 0.0
    for _ in range(12):
        h = 10 ** rng.uniform(-3, -1)
        ys = rng.uniform(-h, h, (40, 3))
        qs = rng.uniform(-1, 1, 40)
        s = math.sqrt(3.0) * h
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        d = u * s * (2.0 + 10 ** rng.uniform(-2, 2))
        M, mhat = {}, {}
        for k in monos(p):
            M[k] = float(np.sum(qs * ys[:, 0] ** k[0] * ys[:, 1] ** k[1] * ys[:, 2] ** k[2]))
            mhat[k] = M[k] / s ** sum(k)
        a = ora.taylor_coeffs(p + 1, d[0], d[1], d[2], float(d @ d) + delta2, kappa)
        ref = sum(M[k] * a[k] for k in monos(p))
        got = far_q(qcoef(mhat, p), p, d, s, kappa, delta2)
        scale = float(np.sum(np.abs(qs))) / math.sqrt(float(d @ d) + delta2)
        worst = max(worst, abs(got - ref) / scale)
    assert worst < 1e-13


def test_plant_dropped_term_fails(ora):
    """A Q with one j-term dropped must fail the same comparison."""
    p, kappa, delta2 = 6, 0.5, 0.0
    rng = np.random.default_rng(7)
    ys = rng.uniform(-0.01, 0.01, (30, 3))
    qs = rng.uniform(-1, 1, 30)
    s = math.sqrt(3.0) * 0.01
    d = np.array([0.05, -0.03, 0.02])
    M = {k: float(np.sum(qs * ys[:, 0] ** k[0] * ys[:, 1] ** k[1] * ys[:, 2] ** k[2]))
         for k in monos(p)}
    mhat = {k: v / s ** sum(k) for k, v in M.items()}
    a = ora.taylor_coeffs(p + 1, d[0], d[1], d[2], float(d @ d), kappa)
    ref = sum(M[k] * a[k] for k in monos(p))
    q = qcoef(mhat, p)
    q[(3, (1, 0, 0))] -= mhat[(3, 0, 0)] / 2  # drop j = e_x from Q_{3,(1,0,0)}
    got = far_q(q, p, d, s, kappa, delta2)
    assert abs(got - ref) / abs(ref) > 1e-9


def test_nqcoef_matches_consumption_order():
    import bench
    for nt in range(1, 14):
        p = nt - 1
        n = sum((p - sum(l)) // 2 + 1 for l in monos(p))
        assert bench.nqcoef(nt) == n


def test_sweep_takes_one_fma_per_coefficient():
    """bench.py prices a far pair at 2 NQ flops: the ratio-Horner sweep of
    far_q (t descending, then l0, l1, l2) takes NQ - 1 multiply-adds (every
    coefficient but each t's first, plus the joins of rows, planes and t
    levels) and one multiplication g_t * phi per t."""
    import bench
    for nt in range(1, 14):
        p = nt - 1
        fma = mul = 0
        for t in range(p // 2, -1, -1):
            L = p - 2 * t
            for l0 in range(L, -1, -1):
                for l1 in range(L - l0, -1, -1):
                    fma += L - l0 - l1          # l2 steps after the row's first
                    fma += l1 != L - l0         # row joined to the plane
                fma += l0 != L                  # plane joined to phi
            mul += 1                            # g_t * phi
            fma += t != p // 2                  # S * z + g_t phi
        assert fma == bench.nqcoef(nt) - 1, nt
        assert mul == p // 2 + 1


def family_weight(l):
    """w_l = prod_{j<=l} (j - 1/2) / j (the regularized family's weights)."""
    w = 1.0
    for j in range(1, l + 1):
        w *= (j - 0.5) / j
    return w


def field_qcoef(mhat, p):
    """The field kernels' per-node coefficients: E_e = sum_k (k_e + 1) M_k a_{k+e}
    is a Taylor sum of degree p + 1 with moments M^(e)_k' = k'_e M_{k'-e}, so
    Q^(e) is qcoef of those, scaled by s (Mhat^(e)_k' = k'_e Mhat_{k'-e} / s; the
    kernel applies the 1/s per pair)."""
    out = []
    for e in range(3):
        me = {}
        for k in monos(p + 1):
            if k[e] >= 1:
                km = list(k)
                km[e] -= 1
                me[k] = k[e] * mhat[tuple(km)]
            else:
                me[k] = 0.0
        out.append(qcoef(me, p + 1))
    return out


def field_q(qs, p, d, s, delta2, korder):
    """The field kernels' form: for each component the ratio-Horner sweep of
    far_q at degree p + 1, with ghat_m = R^(2m+1) (R^-1 d/dR)^m G_L of the
    regularized family G_L = sum_{l<=L} w_l delta^2l R^-(2l+1):
        ghat_m = (-1)^m sum_l w_l (delta^2/R^2)^l prod_{i<m} (2l + 2i + 1),
    times 1/(R s)."""
    P = p + 1
    R = math.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2 + delta2)
    rho = s / R
    w = [rho * d[e] / R for e in range(3)]
    z = rho * rho
    t_l = [family_weight(l) * (delta2 / (R * R)) ** l for l in range(korder + 1)]
    gh = []
    for m in range(P + 1):
        gh.append((-1) ** m * sum(t_l))
        t_l = [t * (2 * l + 2 * m + 1) for l, t in enumerate(t_l)]
    r = [gh[m + 1] / gh[m] for m in range(P)]
    E = []
    for q in qs:
        S = 0.0
        for t in range(P // 2, -1, -1):
            L = P - 2 * t
            phi = 0.0
            for l0 in range(L, -1, -1):
                a = 0.0
                for l1 in range(L - l0, -1, -1):
                    b = 0.0
                    for l2 in range(L - l0 - l1, -1, -1):
                        qv = q[(l0 + l1 + l2 + t, (l0, l1, l2))]
                        b = qv if l2 == L - l0 - l1 else b * r[l0 + l1 + l2 + t] * w[2] + qv
                    a = b if l1 == L - l0 else a * r[l0 + l1 + t] * w[1] + b
                phi = a if l0 == L else phi * r[l0 + t] * w[0] + a
            S = gh[t] * phi if t == P // 2 else S * z + gh[t] * phi
        E.append(S / (R * s))
    return E


@pytest.mark.parametrize("p,delta2,korder", [(0, 0.0, 0), (4, 1e-4, 0), (7, 0.01, 0),
                                             (8, 0.01, 1), (6, 0.02, 3), (11, 1e-4, 2)])
def test_field_radial_form_equals_taylor_sum(ora, p, delta2, korder):
    """The field form against E_e = sum_l w_l delta^2l sum_k (k_e+1) M_k A^(l)_{k+e}
    with the oracle's coefficients of R^-(2l+1) (ora_taylor_coeffs for l = 0,
    ora_powinv_coeffs above), the oracle's treeforce far term."""
    rng = np.random.default_rng(200 + p + korder)
    worst = 0.0
    for _ in range(10):
        h = 10 ** rng.uniform(-2, -0.5)
        ys = rng.uniform(-h, h, (40, 3))
        qsrc = rng.uniform(-1, 1, 40)
        s = math.sqrt(3.0) * h
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        d = u * s * (2.0 + 10 ** rng.uniform(-1, 1.5))
        M = {k: float(np.sum(qsrc * ys[:, 0] ** k[0] * ys[:, 1] ** k[1] * ys[:, 2] ** k[2]))
             for k in monos(p)}
        mhat = {k: v / s ** sum(k) for k, v in M.items()}
        R2 = float(d @ d) + delta2
        ref = np.zeros(3)
        for l in range(korder + 1):
            a = (ora.taylor_coeffs(p + 2, d[0], d[1], d[2], R2, 0.0) if l == 0 else
                 ora.powinv_coeffs(p + 2, l, d[0], d[1], d[2], R2))
            c = family_weight(l) * delta2 ** l
            for k in monos(p):
                for e in range(3):
                    ke = list(k)
                    ke[e] += 1
                    ref[e] += c * (k[e] + 1) * M[k] * a[tuple(ke)]
        got = np.array(field_q(field_qcoef(mhat, p), p, d, s, delta2, korder))
        scale = float(np.sum(np.abs(qsrc))) / R2
        worst = max(worst, float(np.max(np.abs(got - ref))) / scale)
    assert worst < 1e-12
