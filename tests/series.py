"""Truncated multivariate power series in (y1, y2, y3) -- an independent
evaluation of the Taylor coefficients of G(x - y) = exp(-kappa R)/R,
R^2 = |x - y|^2 + delta2, used to pin the recurrences."""
from math import comb, exp, factorial, sqrt

import numpy as np


def _gbinom(a, m):
    r = 1.0
    for i in range(m):
        r *= (a - i) / (i + 1)
    return r


class Series:
    def __init__(self, P, c=None):
        self.P = P
        self.c = np.zeros((P + 1,) * 3) if c is None else c

    def mask(self):
        i, j, k = np.indices((self.P + 1,) * 3)
        return (i + j + k) <= self.P

    def __add__(self, o):
        return Series(self.P, self.c + o.c)

    def scale(self, s):
        return Series(self.P, self.c * s)

    def __mul__(self, o):
        P = self.P
        out = np.zeros_like(self.c)
        for (i, j, k) in zip(*np.nonzero(self.c)):
            v = self.c[i, j, k]
            out[i:, j:, k:] += v * o.c[:P + 1 - i, :P + 1 - j, :P + 1 - k]
        return Series(P, out * self.mask())

    @staticmethod
    def const(P, v):
        s = Series(P)
        s.c[0, 0, 0] = v
        return s


def compose(w, coeffs):
    """sum_m coeffs[m] w^m for a series w with zero constant term."""
    P = w.P
    out = Series.const(P, coeffs[0])
    p = Series.const(P, 1.0)
    for m in range(1, P + 1):
        p = p * w
        out = out + p.scale(coeffs[m])
    return out


def taylor_coeffs(P, x, kappa, delta2):
    """a[i,j,k] = coefficient of y1^i y2^j y3^k in exp(-kappa R)/R, degree <= P."""
    R02 = x[0] ** 2 + x[1] ** 2 + x[2] ** 2 + delta2
    R0 = sqrt(R02)
    w = Series(P)
    for e in range(3):
        idx = [0, 0, 0]
        idx[e] = 1
        w.c[tuple(idx)] = -2.0 * x[e] / R02
        idx[e] = 2
        w.c[tuple(idx)] = 1.0 / R02
    s = compose(w, [0.0] + [_gbinom(0.5, m) for m in range(1, P + 1)])   # sqrt(1+w) - 1
    g = compose(s.scale(-kappa * R0), [1.0 / factorial(m) for m in range(P + 1)])
    h = compose(w, [_gbinom(-0.5, m) for m in range(P + 1)])             # (1+w)^(-1/2)
    return (h * g).scale(exp(-kappa * R0) / R0).c


def powinv_coeffs(P, x, l, delta2):
    """a[i,j,k] = coefficient of y1^i y2^j y3^k in R^-(2l+1), degree <= P."""
    R02 = x[0] ** 2 + x[1] ** 2 + x[2] ** 2 + delta2
    w = Series(P)
    for e in range(3):
        idx = [0, 0, 0]
        idx[e] = 1
        w.c[tuple(idx)] = -2.0 * x[e] / R02
        idx[e] = 2
        w.c[tuple(idx)] = 1.0 / R02
    s = -(2 * l + 1) / 2.0
    h = compose(w, [_gbinom(s, m) for m in range(P + 1)])                # (1+w)^s
    return h.scale(R02 ** s).c
