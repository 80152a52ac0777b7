"""Sphere boundary-integral solve on the treecode (matrix-free GMRES).

Second-kind screened-Poisson equation on the unit sphere, collocated at the
centroids of an icosahedral triangulation with triangle-area weights:

    sigma(x)/2 + 1/(4 pi) \\int_S G(|x - y|) sigma(y) dS(y) = f(x),
    G(r) = exp(-kappa R)/R,  R = sqrt(r^2 + delta^2),

discretised as  sigma_i/2 + sum_j w_j G(x_i - x_j) sigma_j = f_i,  w_j =
area_j/(4 pi), the j = i term being w_i G(0) (finite thanks to delta).  Every
matrix-vector product is one treecode evaluation (bit_solve_second_kind,
FP64 GMRES).

The right-hand side is a seeded combination of zonal spherical harmonics
P_l(x . u), l <= lmax, around random axes u.  G depends on x . y only, so by the
Funk-Hecke theorem P_l(x . u) is an eigenfunction of the integral operator with
eigenvalue lambda_l = 1/2 + (1/(4 pi)) 2 pi \\int_{-1}^{1} G(sqrt(2 - 2t)) P_l(t) dt;
the exact solution of the continuous equation is therefore
sigma = sum a P_l(x . u) / lambda_l, which pins the solve up to the
discretisation error.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
from numpy.polynomial.legendre import leggauss, legval


def icosphere(level: int):
    """Vertices [nv,3] on the unit sphere and faces [20*4^level, 3]."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 dtype=np.float64)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    for _ in range(level):
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * len(V) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a, b = uniq // len(V), uniq % len(V)
        mid = V[a] + V[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = len(V) + inv.reshape(3, -1)  # midpoints of edges 01, 12, 20 per face
        V = np.concatenate([V, mid])
        m01, m12, m20 = m[0], m[1], m[2]
        F = np.concatenate([np.stack([F[:, 0], m01, m20], 1), np.stack([F[:, 1], m12, m01], 1),
                            np.stack([F[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return V, F


def collocation(V, F):
    """Centroids projected to the sphere, flat-triangle areas, mean edge length."""
    A, B, C = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    c = (A + B + C) / 3.0
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    area = 0.5 * np.linalg.norm(np.cross(B - A, C - A), axis=1)
    edges = np.concatenate([np.linalg.norm(B - A, axis=1), np.linalg.norm(C - B, axis=1),
                            np.linalg.norm(A - C, axis=1)])
    return c, area, float(edges.mean())


def kernel(r, kappa, delta2):
    R = np.sqrt(r * r + delta2)
    return np.exp(-kappa * R) / R


def operator_eigenvalues(lmax: int, kappa: float, delta2: float, diag: float = 0.5):
    """lambda_l = diag + (1/2) \\int_{-1}^{1} G(sqrt(2-2t)) P_l(t) dt, l = 0..lmax,
    by Gauss-Legendre on geometrically graded panels in u = 1 - t (the kernel
    is sharply peaked at t = 1 when delta is small)."""
    x, w = leggauss(40)
    edges = [0.0]
    h = max(delta2, 1e-12)
    while edges[-1] < 2.0:
        edges.append(min(2.0, max(h, edges[-1] * 2.0)))
        h = edges[-1]
    lam = np.zeros(lmax + 1)
    for a, b in zip(edges[:-1], edges[1:]):
        u = 0.5 * (b - a) * x + 0.5 * (b + a)
        wu = 0.5 * (b - a) * w
        t = 1.0 - u
        g = kernel(np.sqrt(2.0 * u), kappa, delta2)
        for l in range(lmax + 1):
            c = np.zeros(l + 1)
            c[l] = 1.0
            lam[l] += np.sum(wu * g * legval(t, c))
    return diag + 0.5 * lam


@dataclass
class Rhs:
    axes: np.ndarray     # [k, 3]
    degree: np.ndarray   # [k]
    coef: np.ndarray     # [k]

    def __call__(self, pts):
        out = np.zeros(len(pts))
        for u, l, a in zip(self.axes, self.degree, self.coef):
            c = np.zeros(l + 1)
            c[l] = 1.0
            out += a * legval(pts @ u, c)
        return out

    def solution(self, pts, lam):
        out = np.zeros(len(pts))
        for u, l, a in zip(self.axes, self.degree, self.coef):
            c = np.zeros(l + 1)
            c[l] = 1.0
            out += a / lam[l] * legval(pts @ u, c)
        return out


def random_rhs(lmax: int, seed: int, per_degree: int = 2) -> Rhs:
    rng = np.random.default_rng(seed)
    deg = np.repeat(np.arange(lmax + 1), per_degree)
    ax = rng.standard_normal((len(deg), 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    return Rhs(ax, deg, rng.uniform(-1.0, 1.0, len(deg)))


def solve_sphere(ctx, level: int = 7, kappa: float = 1.0, nterms: int = 7, acc: float = 0.6,
                 restart: int = 50, tol: float = 1e-6, maxit: int = 500, lmax: int = 8,
                 seed: int = 17, maxmemb: int = 500, precision: str = "fp64"):
    """Builds the level-`level` problem, solves it and reports errors/timings."""
    import bit_amd
    V, F = icosphere(level)
    pts, area, h = collocation(V, F)
    d2 = h * h
    w = area / (4.0 * np.pi)
    rhs = random_rhs(lmax, seed)
    f = rhs(pts)
    t0 = time.perf_counter()
    s = bit_amd.System(ctx, pts[:, 0], pts[:, 1], pts[:, 2], w, maxmemb)
    self_coef = float(np.exp(-kappa * h) / h)  # G(0) with R = delta
    t1 = time.perf_counter()
    sigma, it, rr = s.solve_second_kind(w, f, 0.5, self_coef, None, restart, tol, maxit, nterms,
                                        acc, kappa, d2, None, precision)
    t2 = time.perf_counter()
    lam = operator_eigenvalues(lmax, kappa, d2)
    exact = rhs.solution(pts, lam)
    err = float(np.linalg.norm(sigma - exact) / np.linalg.norm(exact))
    s.close()
    return dict(level=level, n=len(pts), delta=h, iters=it, relres=rr, err_vs_modal=err,
                setup_s=t1 - t0, solve_s=t2 - t1, sigma=sigma, points=pts, weights=w, f=f,
                self_coef=self_coef)
