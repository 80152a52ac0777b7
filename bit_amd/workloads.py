"""Seeded synthetic particle sets of the benchmark configurations.

All generators use numpy's PCG64 with an explicit seed, so a (kind, n, seed)
triple names one particle set everywhere (tests, bench, fixtures).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Particles:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    q: np.ndarray

    @property
    def n(self) -> int:
        return len(self.x)


def uniform_cube(n: int, seed: int, lo: float = -0.5, hi: float = 0.5) -> Particles:
    """i.i.d. uniform positions in [lo,hi]^3, charges uniform in [-1,1]."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(lo, hi, size=(n, 3))
    q = rng.uniform(-1.0, 1.0, size=n)
    return Particles(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q)


def gaussian_blobs(n: int, seed: int, nblobs: int = 16, sigma=(0.01, 0.05)) -> Particles:
    """Mixture of isotropic Gaussian blobs: centres uniform in [-0.5,0.5]^3,
    sigma uniform in `sigma`, blob chosen uniformly per particle; charges
    uniform in [-1,1]."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-0.5, 0.5, size=(nblobs, 3))
    sig = rng.uniform(sigma[0], sigma[1], size=nblobs)
    pick = rng.integers(0, nblobs, size=n)
    p = centres[pick] + rng.standard_normal(size=(n, 3)) * sig[pick, None]
    q = rng.uniform(-1.0, 1.0, size=n)
    return Particles(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q)


def anisotropic_box(n: int, seed: int, extent=(1.0, 0.6, 0.35)) -> Particles:
    """Uniform positions in an elongated box: trees with y- and z-splits after
    x-splits (exercises treemake's index quirk, tree3d.cpp:1458-1477)."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.0, 1.0, size=(n, 3)) * np.asarray(extent)
    q = rng.uniform(-1.0, 1.0, size=n)
    return Particles(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q)


def sample_targets(n: int, k: int, seed: int) -> np.ndarray:
    """k distinct particle indices in [0, n), sorted (seeded)."""
    rng = np.random.default_rng(seed)
    k = min(k, n)
    return np.sort(rng.choice(n, size=k, replace=False)).astype(np.int32)


# The benchmark configurations (name -> generator and treecode parameters).
CONFIGS = {
    # Coulomb 1/r, delta=0, kappa=0, theta(acc)=0.5, p=6 (nterms 7), N0=500, FP64
    "uniform10k_coulomb": dict(kind="uniform", n=10_000, seed=1, nterms=7, acc=0.5, kappa=0.0,
                               delta=0.0, maxmemb=500),
    # regularized Coulomb delta=0.01, acc=0.5, p=8, N0=1000
    "uniform1m_regcoulomb": dict(kind="uniform", n=1_000_000, seed=2, nterms=9, acc=0.5,
                                 kappa=0.0, delta=0.01, maxmemb=1000),
    # 16 Gaussian blobs, screened kappa=0.5, delta=0.005, acc=0.7, p=8, N0=2000
    "blobs4m_screened": dict(kind="blobs", n=4_000_000, seed=3, nterms=9, acc=0.7, kappa=0.5,
                             delta=0.005, maxmemb=2000),
}


def make(kind: str, n: int, seed: int) -> Particles:
    if kind == "uniform":
        return uniform_cube(n, seed)
    if kind == "blobs":
        return gaussian_blobs(n, seed)
    if kind == "aniso":
        return anisotropic_box(n, seed)
    raise ValueError(kind)
