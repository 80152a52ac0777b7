"""Benchmark: treecode potential evaluation at N=4M (16 Gaussian blobs, screened).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp32|fp64]
                    [--nterms P1] [--acc THETA] [--maxmemb N0] [--no-cpu-baseline]

One step = one whole evaluation of the potential at all N particles of one
system whose coordinates and charges are already resident in HBM: tree build on
the device (bit-exact with the reference treemake), FP64 moments, batch
interaction lists, far field and near field, result left in HBM.

Prints ONE JSON line.  Extra keys: `accuracy` (relative L2 error of the last
step against an FP64 GPU direct sum on 10,000 seeded targets), `breakdown`
(ms per stage of the last step), `fp64_ms_per_step` (one FP64 step, untimed
for the metric), `configs` (the workload set's other configurations, timed),
`roofline` (dominant kernel), `serving` (p50/p99 request
latency and throughput of a seeded Poisson request stream served with
continuous batching) and `cpu_baseline` (the oracle's OpenMP treecode on the
host cores, extrapolated from a bounded target sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32 = 157.3e12  # MI355X vector FP32, FLOP/s
PEAK_FP64 = 78.6e12   # MI355X vector FP64, FLOP/s

# HBM bytes per launch of the dominant kernel, measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE on this exact configuration (one pass per counter, each scaled by
# its reading for a 1 GiB device-to-device copy in the same process: FETCH x2,
# WRITE x1; the summary is kept under profiles/).  Counters cannot be read from
# inside the timed run, so the figure is reported only when the configuration
# matches the profiled one.
MEASURED_TRAFFIC = {
    # (kernel, n, seed, nterms, acc, maxmemb, kappa, delta, precision): bytes per launch
    # (profiles/pmc_traffic_blobs4m_all_kernels.md)
    ("far", 4_000_000, 3, 12, 0.255, 500, 0.5, 0.005, "fp32"): 5.041481e9,
    ("near", 4_000_000, 3, 12, 0.255, 500, 0.5, 0.005, "fp32"): 1.366242e9,
}
# Executed FP32 lane FLOPs per launch: 64 x rocprofv3's SQ_INSTS_VALU_FLOPS_FP32 (a
# per-wave count, 2 FMA + ADD + MUL + TRANS) on this exact configuration
# (profiles/pmc_flops_k_farq_k_near32_blobs4m.md); masked lanes are included.
MEASURED_EXEC_FLOPS = {
    ("far", 4_000_000, 3, 12, 0.255, 500, 0.5, 0.005, "fp32"): 64 * 2.914759e10,
    ("near", 4_000_000, 3, 12, 0.255, 500, 0.5, 0.005, "fp32"): 64 * 1.303061e10,
}


def nqcoef(nterms: int) -> int:
    """Coefficients per cluster of the FP32 far kernel's radial-derivative form."""
    p = nterms - 1
    return sum((d + 1) * (d + 2) // 2 * ((p - d) // 2 + 1) for d in range(p + 1))


def flops(stats: dict, nterms: int, screened: bool, precision: str) -> tuple[float, float]:
    """Algorithmic FLOPs of the near and far kernels (FMA = 2, transcendental = 1).
    Near, per pair: 16 screened, 12 Coulomb (3 differences, R^2 by FMAs with
    delta^2, rsqrt, the FMA into the sum; screened adds R, -kappa R, exp, q*e).
    Far, per (cluster, target) pair, both precisions (radial-derivative form):
    the ratio-Horner sweep takes one FMA per cluster coefficient, 2 NQ(p); its
    O(p) per-pair set-up is not counted."""
    fn = 16.0 if screened else 12.0
    ff = 2.0 * nqcoef(nterms)
    return stats["near_pairs"] * fn, stats["far_pairs"] * ff


def log(msg: str):
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def cpu_baseline(P, nterms, acc, kappa, d2, maxmemb, sample: int, seed: int = 5):
    """The oracle (FP64 C restatement of the reference treecode, OpenMP over
    targets) on the host: tree + moments timed in full, the per-target
    traversal timed on `sample` seeded targets and extrapolated to N."""
    from bit_amd import workloads
    from oracle.pyoracle import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    O = Oracle()
    t0 = time.perf_counter()
    t = O.tree(P.x, P.y, P.z, maxmemb)
    t1 = time.perf_counter()
    m = t.moments(P.q, nterms)
    t2 = time.perf_counter()
    tidx = workloads.sample_targets(P.n, sample, seed=seed)
    t.potential(P.q, m, nterms, acc, kappa, d2, d2, tidx, nthreads=threads)
    t3 = time.perf_counter()
    per_target = (t3 - t2) / len(tidx)
    total = (t1 - t0) + (t2 - t1) + P.n * per_target
    # the workload set's other potential configurations on the same host cores:
    # 10,000 uniform Coulomb in full, 1M uniform regularized with the traversal
    # timed on 5,000 seeded targets (tree and moments in full)
    configs = {}
    for name, n, seed, mm, nt, d2, samp in (("uniform10k_coulomb", 10_000, 1, 500, 7, 0.0, None),
                                           ("uniform1m_regularized", 1_000_000, 2, 1000, 9,
                                            1e-4, 5000)):
        Q = workloads.uniform_cube(n, seed=seed)
        a0 = time.perf_counter()
        tq = O.tree(Q.x, Q.y, Q.z, mm)
        mq = tq.moments(Q.q, nt)
        a1 = time.perf_counter()
        ti = (np.arange(n, dtype=np.int32) if samp is None else
              workloads.sample_targets(n, samp, seed=seed + 10))
        tq.potential(Q.q, mq, nt, 0.5, 0.0, d2, d2, ti, nthreads=threads)
        a2 = time.perf_counter()
        tot = (a1 - a0) + n * (a2 - a1) / len(ti)
        configs[name] = {"value": n / tot, "unit": "targets/s", "ms_per_evaluation": 1e3 * tot,
                         "sample": ("all targets" if samp is None else
                                    f"traversal on {samp} seeded targets, extrapolated")}
    return {
        "value": P.n / total, "unit": "targets/s", "cores": threads, "kind": "port",
        "configs": configs,
        "cpu_model": cpu_model(), "nproc": os.cpu_count(),
        "sample": f"tree+moments of all {P.n} particles timed in full ({t1 - t0:.2f}s + "
                  f"{t2 - t1:.2f}s); traversal timed on {len(tidx)} seeded targets "
                  f"({t3 - t2:.2f}s on {threads} threads) and extrapolated to N",
        # oracle/time_vs_ref.py: the port against the compiled reference, one thread,
        # 1M uniform, kappa 0.5, nterms 9, acc 0.5, MAC shift 0.01
        "validation": "port 1.109 ms/target vs reference 1.035 ms/target (port/reference "
                      "time 1.07), potentials equal to 7.3e-15 (build container, Xeon 2.7 GHz)",
    }


def field_timings(ctx, log):
    """The field path (treeforce, regularized Coulomb with DEL = 0.01 and the
    MAC's DEL shift) timed as whole evaluations (tree, moments, lists, far,
    near) from HBM-resident inputs: tests/treecode_mp.cpp's size (16,000
    uniform particles, maxmemb 100, nterms 5) and 1M uniform particles
    (maxmemb 1000, nterms 9), acc 0.5, FP64 and FP32; the FP32 field's
    relative L2 distance to the FP64 one is reported."""
    import bit_amd
    from bit_amd import workloads
    out = {}
    for name, n, mm, nt in (("uniform16k", 16_000, 100, 5), ("uniform1m", 1_000_000, 1000, 9)):
        P = workloads.uniform_cube(n, seed=21)
        dev = [bit_amd.DeviceArray.from_numpy(ctx, a) for a in (P.x, P.y, P.z, P.q)]
        rec = {"n": n, "maxmemb": mm, "nterms": nt, "acc": 0.5, "delta2": 0.01}
        E = {}
        for prec in ("fp64", "fp32"):
            def one():
                s = bit_amd.System(ctx, *dev, mm)
                e = s.field(nt, 0.5, 0.01, 0.01, precision=prec)
                st = s.stats()
                s.close()
                return e, st
            one()
            ctx.sync()
            t0 = time.perf_counter()
            reps = 3
            for _ in range(reps):
                E[prec], st = one()
            ms = 1e3 * (time.perf_counter() - t0) / reps
            rec[prec] = {"ms_per_evaluation": ms, "targets_per_s": n / (ms * 1e-3),
                         "ms_far": st["ms_far"], "ms_near": st["ms_near"]}
        rec["fp32_rel_l2_vs_fp64"] = float(np.linalg.norm(E["fp32"] - E["fp64"]) /
                                           np.linalg.norm(E["fp64"]))
        for d in dev:
            d.free()
        log(f"[bench] field {name}: fp64 {rec['fp64']['ms_per_evaluation']:.2f} ms, fp32 "
            f"{rec['fp32']['ms_per_evaluation']:.2f} ms (fp32 vs fp64 {rec['fp32_rel_l2_vs_fp64']:.1e})")
        out[name] = rec
    return out


def config_timings(ctx, log):
    """The other configurations of the workload set, each timed as whole
    evaluations from inputs already in HBM (tree, moments, lists, near and far
    field; one untimed evaluation first, then the mean of three):
    * uniform10k_coulomb: 10,000 uniform particles in [-0.5, 0.5]^3, 1/r,
      theta 0.5, p 6, N0 500, FP64; error against the FP64 direct sum at every
      particle;
    * uniform1m_regularized: 1,000,000 uniform particles, 1/sqrt(r^2 + delta^2),
      delta 0.01, theta 0.5, p 8, N0 1000, FP64 and FP32; error against the FP64
      direct sum on 10,000 seeded targets;
    * sphere_bie_level7: the second-kind screened-Poisson solve on the level-7
      icosphere (327,680 unknowns, kappa 1, delta = mean edge, theta 0.6, p 6,
      GMRES(50) to 1e-6) with FP64 and FP32 products; setup (tree build) and solve
      wall times, iterations, residual and distance to the modal solution."""
    import bit_amd
    from bit_amd import bie, workloads
    out = {}
    for name, n, seed, mm, nt, d2, precs, ntgt in (
            ("uniform10k_coulomb", 10_000, 1, 500, 7, 0.0, ("fp64",), None),
            ("uniform1m_regularized", 1_000_000, 2, 1000, 9, 1e-4, ("fp64", "fp32"), 10_000)):
        P = workloads.uniform_cube(n, seed=seed)
        dev = [bit_amd.DeviceArray.from_numpy(ctx, a) for a in (P.x, P.y, P.z, P.q)]
        res = bit_amd.DeviceArray(ctx, n)
        tidx = (np.arange(n, dtype=np.int32) if ntgt is None else
                workloads.sample_targets(n, ntgt, seed=7))
        s = bit_amd.System(ctx, *dev, mm)
        ref = s.direct(0.0, d2, tidx)
        s.close()
        rec = {"n": n, "maxmemb": mm, "nterms": nt, "p": nt - 1, "acc": 0.5, "delta2": d2,
               "targets_checked": int(len(tidx))}
        for prec in precs:
            def one():
                s = bit_amd.System(ctx, *dev, mm)
                s.potential(nt, 0.5, 0.0, d2, None, prec, out=res)
                s.close()
            one()
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(3):
                one()
            ctx.sync()
            ms = 1e3 * (time.perf_counter() - t0) / 3
            err = float(np.linalg.norm(res.numpy()[tidx] - ref) / np.linalg.norm(ref))
            rec[prec] = {"ms_per_evaluation": ms, "targets_per_s": n / (ms * 1e-3),
                         "rel_l2_vs_fp64_direct": err}
            log(f"[bench] {name} {prec}: {ms:.2f} ms per evaluation, err {err:.2e}")
        for d in dev + [res]:
            d.free()
        out[name] = rec
    rec = {"level": 7, "kappa": 1.0, "acc": 0.6, "nterms": 7, "p": 6, "restart": 50,
           "tol": 1e-6}
    for prec in ("fp64", "fp32"):
        r = bie.solve_sphere(ctx, level=7, kappa=1.0, nterms=7, acc=0.6, restart=50, tol=1e-6,
                             lmax=8, seed=17, maxmemb=500, precision=prec)
        rec["unknowns"] = r["n"]
        rec[prec] = {"setup_ms": 1e3 * r["setup_s"], "solve_ms": 1e3 * r["solve_s"],
                     "iterations": r["iters"], "relres": r["relres"],
                     "rel_l2_vs_modal_solution": r["err_vs_modal"]}
        log(f"[bench] sphere BIE level 7 {prec}: setup {1e3 * r['setup_s']:.1f} ms, solve "
            f"{1e3 * r['solve_s']:.1f} ms, {r['iters']} iterations, relres {r['relres']:.2e}")
    out["sphere_bie_level7"] = rec
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=4_000_000)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--nterms", type=int, default=12)
    ap.add_argument("--acc", type=float, default=0.255)
    ap.add_argument("--maxmemb", type=int, default=500)
    ap.add_argument("--kappa", type=float, default=0.5)
    ap.add_argument("--delta", type=float, default=0.005)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--accuracy-targets", type=int, default=10_000)
    ap.add_argument("--cpu-sample", type=int, default=20000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--fp64-point", default="12,0.255,250",
                    help="nterms,acc,maxmemb of the FP64 evaluation compared with the FP32 "
                         "step (the autotuner's FP64 point at 1e-6)")
    ap.add_argument("--no-literal", action="store_true",
                    help="skip the literal-parameter runs (p 8, acc 0.7, N0 2000)")
    ap.add_argument("--no-field", action="store_true",
                    help="skip the field (treeforce) timings at 16k and 1M particles")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other configurations (10k Coulomb, 1M regularized, sphere "
                         "BIE at level 7)")
    ap.add_argument("--serving-requests", type=int, default=2000,
                    help="requests of the serving leg (0 = skip)")
    ap.add_argument("--serving-rate", type=float, default=400.0)
    ap.add_argument("--serving-sweep", default="1600,6400",
                    help="extra arrival rates (req/s) served by the same server, 1000 "
                         "requests each (empty = none)")
    ap.add_argument("--no-serving-baseline", action="store_true",
                    help="skip the one-at-a-time run of the serving stream")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also copy this many MB device-to-device (a plain copy kernel of "
                         "known size, for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the treecode's multi-GPU path (target shards, "
              "one broadcast of sources and moments) is a design, not part of this build; "
              "run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import bit_amd
    from bit_amd import workloads

    d2 = args.delta ** 2
    screened = args.kappa != 0.0
    log(f"[bench] generating {args.n} blob particles (seed {args.seed})")
    P = workloads.gaussian_blobs(args.n, seed=args.seed)
    ctx = bit_amd.Context(0)
    dev = [bit_amd.DeviceArray.from_numpy(ctx, a) for a in (P.x, P.y, P.z, P.q)]
    out = bit_amd.DeviceArray(ctx, args.n)

    if args.copy_calibration_mb > 0:
        m = args.copy_calibration_mb * (1 << 20) // 8
        a_ = bit_amd.DeviceArray(ctx, m)
        b_ = bit_amd.DeviceArray(ctx, m)
        b_.copy_from(a_)
        a_.free()
        b_.free()
        log(f"[bench] copied {args.copy_calibration_mb} MB device-to-device (calibration)")

    def step(precision):
        s = bit_amd.System(ctx, *dev, args.maxmemb)
        s.potential(args.nterms, args.acc, args.kappa, d2, None, precision, out=out)
        st = s.stats()
        s.close()
        return st

    for w in range(args.warmup):
        t0 = time.perf_counter()
        step(args.precision)
        log(f"[bench] warmup {w}: {1e3 * (time.perf_counter() - t0):.1f} ms")
    ctx.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        st = step(args.precision)
    ctx.sync()  # every step already ends synchronised (results resident in HBM)
    t1 = time.perf_counter()
    ms = 1e3 * (t1 - t0) / args.steps
    log(f"[bench] {args.steps} steps: {ms:.1f} ms/step")

    # accuracy of the last step's result against the FP64 direct sum
    tidx = workloads.sample_targets(args.n, args.accuracy_targets, seed=7)
    phi = out.numpy()
    s = bit_amd.System(ctx, *dev, args.maxmemb)
    ref = s.direct(args.kappa, d2, tidx)
    s.close()
    err = float(np.linalg.norm(phi[tidx] - ref) / np.linalg.norm(ref))
    log(f"[bench] rel L2 error vs FP64 direct sum on {len(tidx)} targets: {err:.3e}")

    def timed(nterms, acc, maxmemb, precision, mac_sub=None, steps=3):
        """ms per step (one warm-up, then `steps` whole evaluations) and the rel
        L2 error of the last one against the FP64 direct sum on tidx."""
        def one():
            s = bit_amd.System(ctx, *dev, maxmemb)
            s.potential(nterms, acc, args.kappa, d2, mac_sub, precision, out=out)
            r = s.stats()
            s.close()
            return r
        one()
        ctx.sync()
        a = time.perf_counter()
        for _ in range(steps):
            r = one()
        ctx.sync()
        t = 1e3 * (time.perf_counter() - a) / steps
        e = float(np.linalg.norm(out.numpy()[tidx] - ref) / np.linalg.norm(ref))
        return t, e, r

    fp64 = None
    if not args.no_fp64_compare and args.precision == "fp32":
        # FP64 at the FP32 point, and at the autotuner's FP64 point for the same target
        nt64, acc64, mm64 = args.fp64_point.split(",")
        nt64, acc64, mm64 = int(nt64), float(acc64), int(mm64)
        t_same, e_same, st64 = timed(args.nterms, args.acc, args.maxmemb, "fp64", steps=2)
        t_tuned, e_tuned, _ = timed(nt64, acc64, mm64, "fp64", steps=2)
        fp64 = {"same_point": {"ms_per_step": t_same, "rel_l2_vs_fp64_direct": e_same,
                               "ms_far": st64["ms_far"], "ms_near": st64["ms_near"]},
                "fp64_tuned_point": {"nterms": nt64, "acc": acc64, "maxmemb": mm64,
                                     "ms_per_step": t_tuned, "rel_l2_vs_fp64_direct": e_tuned},
                "fp32_speedup_vs_fp64_tuned": t_tuned / ms}
        log(f"[bench] fp64 at the fp32 point: {t_same:.1f} ms (err {e_same:.2e}); at its own "
            f"tuned point ({nt64},{acc64},{mm64}): {t_tuned:.1f} ms (err {e_tuned:.2e})")
    fp64_ms = fp64["same_point"]["ms_per_step"] if fp64 else None

    literal = None
    if not args.no_literal:
        # the configuration's literal parameters: p 8, acc 0.7, N0 2000, FP32, with
        # the clean MAC (delta^2 subtracted) and with the reference's MAC shift 0.01
        literal = {"nterms": 9, "p": 8, "acc": 0.7, "maxmemb": 2000, "precision": "fp32"}
        for name, sub in (("clean_mac", None), ("reference_mac_shift_0.01", 0.01)):
            t, e, _ = timed(9, 0.7, 2000, "fp32", sub)
            literal[name] = {"ms_per_step": t, "targets_per_s": args.n / (t * 1e-3),
                             "rel_l2_vs_fp64_direct": e}
            log(f"[bench] literal point ({name}): {t:.1f} ms/step, err {e:.2e}")

    f_near, f_far = flops(st, args.nterms, screened, args.precision)
    kern = "far" if st["ms_far"] >= st["ms_near"] else "near"
    kms = st["ms_far"] if kern == "far" else st["ms_near"]
    kfl = f_far if kern == "far" else f_near
    peak = PEAK_FP32 if args.precision == "fp32" else PEAK_FP64
    achieved = kfl / (kms * 1e-3)
    roof = {"bound": "valu", "kernel": (f"k_farq<{args.nterms},{str(screened).lower()}>"
                                        if args.precision == "fp32" else
                                        f"k_farq64<{args.nterms},{str(screened).lower()}>")
            if kern == "far" else
            f"k_near{'32' if args.precision == 'fp32' else '64'}<{str(screened).lower()}>",
            "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": MEASURED_TRAFFIC.get((kern, args.n, args.seed, args.nterms, args.acc,
                                             args.maxmemb, args.kappa, args.delta,
                                             args.precision)),
            "flops_per_launch": kfl, "kernel_ms": kms,
            # per far pair: what is priced (the sweep's FMAs) and what the
            # reference's coefficient recurrence would cost (12 or 24 per Taylor
            # coefficient), which this far kernel does not execute
            "far_flops_per_pair": {"priced": 2 * nqcoef(args.nterms),
                                   "coefficient_recurrence": (24 if screened else 12) *
                                   (args.nterms * (args.nterms + 1) * (args.nterms + 2) // 6)},
            "other_kernel": {"name": "near" if kern == "far" else "far",
                             "tflops": (f_near / (st["ms_near"] * 1e-3) if kern == "far" else
                                        f_far / (st["ms_far"] * 1e-3)) / 1e12}}
    # `frac` prices the far kernel at the sweep it runs (2 NQ(p) per pair); beside it the
    # counter-measured executed FLOPs and the reference recurrence's 24 T(p) per pair,
    # which the kernel does not execute (hence a fraction above 1)
    key = (kern, args.n, args.seed, args.nterms, args.acc, args.maxmemb, args.kappa, args.delta,
           args.precision)
    ex = MEASURED_EXEC_FLOPS.get(key)
    if ex is not None:
        roof["executed"] = {"flops_per_launch": ex, "tflops": ex / (kms * 1e-3) / 1e12,
                            "frac": ex / (kms * 1e-3) / peak,
                            "source": "rocprofv3 SQ_INSTS_VALU_FLOPS_FP32 x 64 lanes"}
    # the FMA issue ceiling of the build's instruction mix (no packed FP32: 78.6
    # TFLOP/s at 2.4 GHz, half the published FP32 peak, which counts packed FMAs),
    # measured live; the executed rate is quoted against it
    c32, c64 = ctx.valu_ceiling()
    ceil = c32 if args.precision == "fp32" else c64
    roof["valu_issue_ceiling"] = {"tflops": ceil, "frac": achieved / (ceil * 1e12)}
    if "executed" in roof:
        roof["valu_issue_ceiling"]["executed_frac"] = roof["executed"]["tflops"] / ceil
    log(f"[bench] FMA issue ceiling {c32:.1f} (fp32) / {c64:.1f} (fp64) TFLOP/s")
    if kern == "far":
        rec = roof["far_flops_per_pair"]["coefficient_recurrence"] * st["far_pairs"]
        roof["recurrence_pricing"] = {"flops_per_launch": rec,
                                      "tflops": rec / (kms * 1e-3) / 1e12,
                                      "frac": rec / (kms * 1e-3) / peak}

    field = None
    if not args.no_field:
        field = field_timings(ctx, log)
    configs = None
    if not args.no_configs:
        configs = config_timings(ctx, log)

    serving = None
    if args.serving_requests > 0:
        from bit_amd import serving as srv_mod
        log(f"[bench] serving leg: {args.serving_requests} requests at {args.serving_rate}/s")
        for d in dev + [out]:
            d.free()
        sweep = [float(v) for v in args.serving_sweep.split(",") if v]
        serving = srv_mod.benchmark(args.serving_requests, args.serving_rate, seed=2024,
                                    builders=4, hbm_budget=16 << 30, warmup=64,
                                    baseline=not args.no_serving_baseline, sweep=sweep,
                                    sweep_requests=1000, log=log)
        serving["workload"] = ("Poisson arrivals; N log-uniform [1e3,2e5]; uniform/blobs; "
                               "delta U[0,0.02], kappa U[0,2], p in {4,6,8}; 20% sphere BIE "
                               "GMRES(50) solves (levels 3-5), Krylov state resident in HBM")

    cpu = None
    if not args.no_cpu_baseline:
        log(f"[bench] CPU baseline (oracle port), sample {args.cpu_sample} targets")
        cpu = cpu_baseline(P, args.nterms, args.acc, args.kappa, d2, args.maxmemb,
                           args.cpu_sample)
        log(f"[bench] CPU baseline {cpu['value']:.3e} targets/s on {cpu['cores']} threads")
        if configs:
            for name, c in cpu["configs"].items():
                if name in configs:
                    configs[name]["cpu_port_targets_per_s"] = c["value"]
                    log(f"[bench] {name} on the host: {c['value']:.3e} targets/s")

    line = {
        "metric": "targets/sec at N=4M, rel L2 err<=1e-6 vs FP64; p50/p99 request latency",
        "value": args.n / (ms * 1e-3),
        "unit": "targets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (seeded 16-blob Gaussian mixture, charges U[-1,1])",
        "config": {"workload": "blobs4m_screened", "n": args.n, "kappa": args.kappa,
                   "delta": args.delta, "nterms": args.nterms, "p": args.nterms - 1,
                   "acc": args.acc, "maxmemb": args.maxmemb, "precision": args.precision,
                   "parallelism": "single GPU"},
        "accuracy": {"rel_l2_vs_fp64_direct": err, "targets": len(tidx)},
        "breakdown": {k: st[k] for k in ("ms_tree", "ms_upload", "ms_moments", "ms_lists",
                                         "ms_far", "ms_near")},
        "interactions": {k: st[k] for k in ("near_pairs", "far_pairs", "far_items",
                                            "near_items", "lane_far", "lane_near_pairs",
                                            "list_passes")},
        "fp64_ms_per_step": fp64_ms,
        "fp64": fp64,
        "literal": literal,
        "field": field,
        "configs": configs,
        "roofline": roof,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
