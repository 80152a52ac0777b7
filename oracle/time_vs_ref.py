"""Per-target time of the oracle's traversal against the reference's own
(oracle/_ref/libbitref.so), both single-threaded, on the same targets of the
same system: the check that the CPU baseline `bench.py` reports (the oracle,
kind "port") runs at the reference's speed.

    OMP_NUM_THREADS=1 python oracle/time_vs_ref.py [--n 1000000] [--targets 2000]

Configuration (the survey's 1M-uniform probe): N uniform in [-0.5,0.5]^3,
shielded_treepot with kappa 0.5, delta 0, nterms 9, acc 0.5, maxmemb 1000, the
reference's MAC shift DEL = 0.01.  The reference call builds its tree and
moments each time, so its per-target time is the difference of two calls with
k1 < k2 targets, divided by k2 - k1.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from bit_amd import workloads  # noqa: E402
from oracle.pyoracle import Oracle, Reference  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--targets", type=int, default=2000)
    a = ap.parse_args()
    nterms, acc, kappa, maxmemb = 9, 0.5, 0.5, 1000
    P = workloads.uniform_cube(a.n, seed=2)
    k1, k2 = a.targets // 10, a.targets
    t_small = workloads.sample_targets(P.n, k1, seed=31)
    t_big = workloads.sample_targets(P.n, k2, seed=32)

    R = Reference()
    c0 = time.perf_counter()
    r1 = R.shielded_phi(P.x, P.y, P.z, P.q, maxmemb, nterms, acc, kappa, t_small)
    c1 = time.perf_counter()
    r2 = R.shielded_phi(P.x, P.y, P.z, P.q, maxmemb, nterms, acc, kappa, t_big)
    c2 = time.perf_counter()
    ref_per_target = ((c2 - c1) - (c1 - c0)) / (k2 - k1)

    O = Oracle()
    o0 = time.perf_counter()
    t = O.tree(P.x, P.y, P.z, maxmemb)
    o1 = time.perf_counter()
    m = t.moments(P.q, nterms)
    o2 = time.perf_counter()
    phi = t.potential(P.q, m, nterms, acc, kappa, 0.0, 0.01, t_big, nthreads=1)
    o3 = time.perf_counter()
    ora_per_target = (o3 - o2) / k2
    err = float(np.max(np.abs(phi - r2)) / np.max(np.abs(r2)))
    print(json.dumps({
        "config": {"n": a.n, "kappa": kappa, "nterms": nterms, "acc": acc, "maxmemb": maxmemb,
                   "mac_shift": 0.01, "targets": [k1, k2], "threads": 1},
        "reference_ms_per_target": 1e3 * ref_per_target,
        "oracle_ms_per_target": 1e3 * ora_per_target,
        "oracle_over_reference_time": ora_per_target / ref_per_target,
        "oracle_tree_s": o1 - o0, "oracle_moments_s": o2 - o1,
        "reference_call_s": [c1 - c0, c2 - c1],
        "rel_max_oracle_vs_reference": err,
    }))


if __name__ == "__main__":
    main()
