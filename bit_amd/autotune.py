"""Autotuner for the treecode parameters (p, N0, theta) at a target accuracy.

For a given particle set and kernel (kappa, delta) it measures, on the GPU,
both the error (relative L2 against an FP64 direct sum on a seeded sample of
targets) and the cost (the wall time of one whole evaluation from inputs
resident in HBM, as the benchmark's step runs it: device tree build, moments on
a side stream beside the lists and the near field, far field) of candidate
parameter sets, and returns the cheapest one whose error stays below the
target with a safety margin.  For each (N0, nterms) the MAC parameter is
scanned from the largest (cheapest) value down and the scan stops at the
first value that meets the target, since the error decreases monotonically as
theta shrinks.  Targets are batched 64 per wavefront, so the batch size is
fixed by the hardware and not searched.

    python -m bit_amd.autotune --n 4000000 --kappa 0.5 --delta 0.005
"""
from __future__ import annotations

import json
import time

import numpy as np

import bit_amd
from bit_amd import workloads

DEFAULT_N0 = (350, 500, 750, 1000, 2000)
DEFAULT_NTERMS = (10, 11, 12, 13)
DEFAULT_ACC = (0.4, 0.35, 0.3, 0.25, 0.2, 0.15)


def autotune(ctx, P, kappa: float, delta: float, target: float = 1e-6, margin: float = 0.8,
             precision: str = "fp32", n0s=DEFAULT_N0, nterms_list=DEFAULT_NTERMS,
             accs=DEFAULT_ACC, sample: int = 4000, seed: int = 11, verbose: bool = False):
    d2 = delta * delta
    dev = [bit_amd.DeviceArray.from_numpy(ctx, a) for a in (P.x, P.y, P.z, P.q)]
    out = bit_amd.DeviceArray(ctx, P.n)
    tidx = workloads.sample_targets(P.n, sample, seed)
    table = []
    best = None
    ref = None
    try:
        for n0 in n0s:
            if ref is None:
                s = bit_amd.System(ctx, *dev, n0)
                ref = s.direct(kappa, d2, tidx)
                s.close()
            for nt in nterms_list:
                for acc in accs:  # largest (cheapest) first
                    ctx.sync()
                    t0 = time.perf_counter()
                    s = bit_amd.System(ctx, *dev, n0)
                    s.potential(nt, acc, kappa, d2, None, precision, out=out)
                    ms = 1e3 * (time.perf_counter() - t0)
                    st = s.stats()
                    s.close()
                    phi = out.numpy()[tidx]
                    err = float(np.linalg.norm(phi - ref) / np.linalg.norm(ref))
                    row = dict(maxmemb=n0, nterms=nt, acc=acc, err=err, ms=ms,
                               far_ms=st["ms_far"], near_ms=st["ms_near"])
                    table.append(row)
                    if verbose:
                        print(json.dumps(row), flush=True)
                    if err <= margin * target:
                        if best is None or ms < best["ms"]:
                            best = row
                        break
    finally:
        for d in dev + [out]:
            d.free()
    return best, table


def main():
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=4_000_000)
    ap.add_argument("--kind", default="blobs")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--kappa", type=float, default=0.5)
    ap.add_argument("--delta", type=float, default=0.005)
    ap.add_argument("--target", type=float, default=1e-6)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--n0s", default=",".join(map(str, DEFAULT_N0)),
                    help="comma-separated leaf sizes")
    ap.add_argument("--nterms", default=",".join(map(str, DEFAULT_NTERMS)),
                    help="comma-separated expansion sizes (p + 1)")
    ap.add_argument("--accs", default=",".join(map(str, DEFAULT_ACC)),
                    help="comma-separated MAC parameters, largest first")
    a = ap.parse_args()
    P = workloads.make(a.kind, a.n, a.seed)
    ctx = bit_amd.Context(0)
    t0 = time.perf_counter()
    best, table = autotune(ctx, P, a.kappa, a.delta, a.target, precision=a.precision,
                           n0s=[int(v) for v in a.n0s.split(",")],
                           nterms_list=[int(v) for v in a.nterms.split(",")],
                           accs=[float(v) for v in a.accs.split(",")], verbose=True)
    print(json.dumps({"best": best, "candidates": len(table),
                      "tuning_s": time.perf_counter() - t0}), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
