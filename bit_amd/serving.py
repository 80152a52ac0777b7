"""Continuous-batching request server for independent treecode requests.

Requests (potential evaluations of independent particle systems, and sphere
boundary-integral GMRES solves) arrive over time.  The server:

* admits arrived requests in order while their estimated device memory fits
  the HBM budget;
* builds each admitted request's tree, moments and interaction lists on a pool
  of builder threads (each with its own context/stream, a few host threads for
  the tree);
* runs a GPU loop of rounds (bit_serve_round): each round takes every ready
  evaluation and one step of every running GMRES solve -- its matrix-vector
  product goes into the round's persistent launch with the evaluations (one
  launch per precision: device-sized work-groups take (system, batch) items
  from a counter, all nterms and kernels in the same launch), and its Arnoldi
  step (two Gram-Schmidt passes, the norm, the updates) is batched with the
  other solves' -- so requests arriving while a round runs join the next one,
  and a solve occupies a slot for as many rounds as it needs products;
* keeps each solve's tree, lists and Krylov state (basis, residual, iterate)
  resident in HBM from admission to completion; only the Hessenberg column and
  the Givens rotations are on the host.

Every target's sums and every solve's vector arithmetic are formed the same
way alone and in a mix, so results do not depend on what else was on the GPU
(tests/test_gpu_serving.py).
"""
from __future__ import annotations

import functools
import math
import queue
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

import bit_amd
from bit_amd import bie, workloads

BYTES_PER_PARTICLE = 320  # sources, FP32 copy, lists, outputs: calibrated upper estimate
BYTES_FIXED = 2 << 20


@dataclass
class Request:
    rid: int
    arrival: float        # seconds from the start of the stream
    kind: str             # "eval" or "bie"
    n: int
    seed: int
    layout: str = "uniform"   # eval: "uniform" or "blobs"
    nterms: int = 7
    acc: float = 0.5
    kappa: float = 0.0
    delta: float = 0.0
    maxmemb: int = 256
    level: int = 0            # bie: icosphere level
    tol: float = 1e-6
    # filled by a run
    t_admit: float = 0.0
    t_ready: float = 0.0
    t_done: float = 0.0
    data: tuple | None = None  # (x, y, z, q, w) made ahead by materialize()
    result: np.ndarray | None = None
    iters: int = 0
    relres: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def est_bytes(self) -> int:
        b = BYTES_PER_PARTICLE * self.n + BYTES_FIXED
        if self.kind == "bie":  # resident Krylov basis (restart 50) and work vectors
            b += 8 * (51 + 6) * self.n
        return b

    def params(self, precision="fp32"):
        return bit_amd.make_params(self.nterms, self.acc, self.kappa, self.delta ** 2, None,
                                   precision)


def make_requests(count: int = 2000, rate: float = 400.0, seed: int = 2024,
                  bie_fraction: float = 0.2, nmin: int = 1_000, nmax: int = 200_000):
    """The serving workload: Poisson arrivals at `rate` per second; N log-uniform
    in [nmin, nmax]; uniform or blob positions; delta ~ U[0, 0.02],
    kappa ~ U[0, 2], p in {4, 6, 8}; a fraction of small sphere BIE solves at
    icosphere levels 3-5 (kappa 1, delta = mean edge, theta 0.6, p 6)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=count))
    reqs = []
    for i in range(count):
        if rng.uniform() < bie_fraction:
            level = int(rng.integers(3, 6))
            reqs.append(Request(i, float(t[i]), "bie", 20 * 4 ** level, int(rng.integers(1 << 30)),
                                nterms=7, acc=0.6, kappa=1.0, level=level, maxmemb=128))
        else:
            n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
            reqs.append(Request(i, float(t[i]), "eval", n, int(rng.integers(1 << 30)),
                                layout="uniform" if rng.uniform() < 0.5 else "blobs",
                                nterms=int(rng.choice([5, 7, 9])), acc=0.5,
                                kappa=float(rng.uniform(0.0, 2.0)),
                                delta=float(rng.uniform(0.0, 0.02)),
                                maxmemb=int(min(256, max(32, n // 16)))))
    return reqs


@functools.lru_cache(maxsize=8)
def _sphere(level: int):
    """Collocation points, weights and mean edge of the icosphere (read-only)."""
    V, F = bie.icosphere(level)
    pts, area, h = bie.collocation(V, F)
    return pts, area / (4.0 * np.pi), h


def _particles(r: Request):
    """The request's data: positions, charges (and, for a solve, the weights);
    a solve's right-hand side, delta and self term go into r.  Made once
    (materialize) before the stream starts, as a client would send it."""
    if r.data is not None:
        return r.data
    if r.kind == "eval":
        P = workloads.make(r.layout, r.n, r.seed)
        return P.x, P.y, P.z, P.q, None
    pts, w, h = _sphere(r.level)
    r.delta = h
    r.extra["f"] = bie.random_rhs(8, r.seed % 100000)(pts)
    r.extra["self"] = float(np.exp(-r.kappa * h) / h)
    return pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy(), w, w


def materialize(reqs, workers: int = 8):
    """Makes every request's data ahead of the stream (the synthetic client's
    work, not the server's)."""
    def one(r):
        r.data = _particles(r)
    with ThreadPoolExecutor(max_workers=workers) as pool:
        list(pool.map(one, reqs))
    return reqs


def solve_alone(ctx, r: Request, precision="fp32"):
    """The request served by itself: one evaluation, or the whole device-resident
    GMRES solve (reference result for the mix tests, and the one-at-a-time
    baseline)."""
    x, y, z, q, w = _particles(r)
    s = bit_amd.System(ctx, x, y, z, q, r.maxmemb)
    try:
        args = (r.nterms, r.acc, r.kappa, r.delta ** 2, None, precision)
        if r.kind == "eval":
            return s.potential(*args)
        sig, it, rr = s.solve_second_kind(w, r.extra["f"], 0.5, r.extra["self"], None, 50,
                                          r.tol, 500, *args)
        r.iters, r.relres = it, rr
        return sig
    finally:
        s.close()


class _Job:
    __slots__ = ("req", "system", "params")

    def __init__(self, req, system, params):
        self.req, self.system, self.params = req, system, params


class Server:
    """Continuous batching on one GPU: builder threads make each admitted
    request's system (tree, moments, lists; for a solve also its resident
    Krylov state), and the GPU loop serves rounds (bit_serve_round): every
    ready evaluation and one step of every running solve in shared launches.
    Four builders by default: at 6,400 requests/s four serve 88-90M targets/s
    against 82M with eight (every builder's device work competes with the
    rounds), and at 400/s the latencies are the same."""

    def __init__(self, device: int = 0, hbm_budget: int = 16 << 30, builders: int = 4,
                 host_threads: int = 2, precision: str = "fp32", max_round_systems: int = 256):
        self.device = device
        self.budget = hbm_budget
        self.builders = builders
        self.host_threads = host_threads
        self.precision = precision
        self.max_round = max_round_systems
        self.gpu_ctx = bit_amd.Context(device)
        self._tls = threading.local()
        self._ctxs = []
        self._lock = threading.Lock()

    def close(self):
        for c in self._ctxs:
            c.close()
        self.gpu_ctx.close()

    def _builder_ctx(self):
        c = getattr(self._tls, "ctx", None)
        if c is None:
            c = bit_amd.Context(self.device)
            c.set_host_threads(self.host_threads)
            with self._lock:
                self._ctxs.append(c)
            self._tls.ctx = c
        return c

    def run(self, requests, realtime: bool = True):
        """Serves `requests` (arrival times relative to the start); returns metrics."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        inbox: "queue.Queue" = queue.Queue()  # ("eval", job) or ("bie", (req, system, krylov))
        inflight = [0]
        cv = threading.Condition()
        pending = list(reversed(reqs))
        finished = [0]
        total = len(reqs)
        t0 = time.perf_counter()
        pool = ThreadPoolExecutor(max_workers=self.builders)
        rounds = []
        build_ms = []

        def now():
            return time.perf_counter() - t0

        def release(r, s, k=None):
            if k is not None:
                k.close()
            s.close()
            r.data = None  # the request's input arrays are no longer needed
            with cv:
                inflight[0] -= r.est_bytes
                finished[0] += 1
                cv.notify_all()

        def fail(r, e):
            r.extra["error"] = repr(e)
            r.t_done = now()
            r.data = None
            with cv:
                inflight[0] -= r.est_bytes
                finished[0] += 1
                cv.notify_all()

        def build(r):
            s = None
            try:
                ctx = self._builder_ctx()
                x, y, z, q, w = _particles(r)
                s = bit_amd.System(ctx, x, y, z, q, r.maxmemb)
                prm = r.params(self.precision)
                if r.kind == "eval":
                    s.prepare(r.nterms, r.acc, r.kappa, r.delta ** 2, None, self.precision)
                    k = None
                else:
                    k = bit_amd.Krylov(s, w, r.extra["f"], 0.5, r.extra["self"], None, 50, r.tol,
                                       500, r.nterms, r.acc, r.kappa, r.delta ** 2, None,
                                       self.precision)
                st = s.stats()
                build_ms.append(max(st["ms_tree"], 0.0) + max(st["ms_moments"], 0.0) +
                                max(st["ms_lists"], 0.0))
            except Exception as e:  # a failed request must not stall the loop
                if s is not None:
                    s.close()
                fail(r, e)
                return
            r.t_ready = now()
            if k is None:
                inbox.put(("eval", _Job(r, s, prm)))
            else:
                inbox.put(("bie", (r, s, k)))

        def admit():
            while pending:
                r = pending[-1]
                if realtime:
                    wait = r.arrival - now()
                    if wait > 0:
                        time.sleep(min(wait, 0.002))
                        continue
                with cv:
                    while inflight[0] > 0 and inflight[0] + r.est_bytes > self.budget:
                        cv.wait(0.01)
                    inflight[0] += r.est_bytes
                pending.pop()
                r.t_admit = now()
                pool.submit(build, r)

        adm = threading.Thread(target=admit, daemon=True)
        adm.start()
        evals, solves = [], []
        while finished[0] < total:
            if not evals and not solves:
                try:
                    kind, item = inbox.get(timeout=0.005)
                except queue.Empty:
                    continue
                (evals if kind == "eval" else solves).append(item)
            while True:
                try:
                    kind, item = inbox.get_nowait()
                except queue.Empty:
                    break
                (evals if kind == "eval" else solves).append(item)
            jobs, evals = evals[:self.max_round], evals[self.max_round:]
            outs, ms = bit_amd.serve_round(self.gpu_ctx, [j.system for j in jobs],
                                           [j.params for j in jobs], [k for _, _, k in solves])
            rounds.append((len(jobs) + len(solves),
                           sum(j.system.n for j in jobs) + sum(s.n for _, s, _ in solves), ms))
            t = now()
            for j, o in zip(jobs, outs):
                j.req.result = o
                j.req.t_done = t
                release(j.req, j.system)
            running = []
            for r, s, k in solves:
                done, it, rr = k.status()
                if not done:
                    running.append((r, s, k))
                    continue
                r.result, r.iters, r.relres = k.solution(), it, rr
                r.t_done = now()
                if done != 1:
                    r.extra["error"] = f"GMRES stopped at maxit (relres {rr:.3e})"
                release(r, s, k)
            solves = running
        adm.join()
        pool.shutdown(wait=True)
        makespan = max(r.t_done for r in reqs) - reqs[0].arrival
        out = summarize(reqs, makespan, rounds)
        # the rounds run one after another on the server's stream: their GPU
        # time over the makespan is the fraction of time the GPU serves rounds;
        # the builders' trees, moments and lists (their own contexts and
        # streams) overlap the rounds and each other, so they are reported apart
        out["gpu_busy_frac"] = out.get("gpu_kernel_s", 0.0) / makespan
        out["build_s"] = float(sum(build_ms) / 1e3)
        return out


def run_one_at_a_time(requests, device: int = 0, precision: str = "fp32",
                      realtime: bool = True):
    """Baseline: the same arrivals served one request at a time (one system,
    one evaluation or one whole solve, on one stream)."""
    ctx = bit_amd.Context(device)
    reqs = sorted(requests, key=lambda r: r.arrival)
    t0 = time.perf_counter()
    for r in reqs:
        if realtime:
            wait = r.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        r.t_admit = r.t_ready = time.perf_counter() - t0
        r.result = solve_alone(ctx, r, precision)
        r.t_done = time.perf_counter() - t0
        r.data = None
    ctx.close()
    makespan = max(r.t_done for r in reqs) - reqs[0].arrival
    return summarize(reqs, makespan, [])


def summarize(reqs, makespan, rounds):
    lat = np.array([r.t_done - r.arrival for r in reqs]) * 1e3
    targets = sum(r.n * (max(r.iters, 1) if r.kind == "bie" else 1) for r in reqs)
    out = {
        "requests": len(reqs),
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "makespan_s": float(makespan),
        "targets_per_s": float(targets / makespan),
        "bie_requests": sum(r.kind == "bie" for r in reqs),
    }
    if rounds:
        out["rounds"] = len(rounds)
        out["mean_systems_per_round"] = float(np.mean([r[0] for r in rounds]))
        out["gpu_kernel_s"] = float(sum(r[2] for r in rounds) / 1e3)
    return out


def benchmark(requests: int = 2000, rate: float = 400.0, seed: int = 2024, builders: int = 4,
              hbm_budget: int = 16 << 30, warmup: int = 64, baseline: bool = True,
              sweep=(), sweep_requests: int = 1000, log=None):
    """The serving stream at `rate` and `sweep_requests` more at each rate of
    `sweep`, each served with continuous batching and (baseline) the same
    arrivals one request at a time."""
    srv = Server(hbm_budget=hbm_budget, builders=builders)
    try:
        if warmup > 0:  # first use of kernel instantiations and builder contexts
            srv.run(materialize(make_requests(warmup, rate, seed + 1)))
        reqs = materialize(make_requests(requests, rate, seed))
        m = srv.run(reqs)
        out = {"requests": requests, "rate_per_s": rate, "seed": seed, "warmup_requests": warmup,
               "server": m, "errors": sum("error" in r.extra for r in reqs)}
        if log:
            log(f"[serving] {requests} requests at {rate}/s: p50 {m['p50_ms']:.1f} ms, p99 "
                f"{m['p99_ms']:.1f} ms, {m['mean_systems_per_round']:.2f} systems/round, "
                f"GPU busy (rounds) {m['gpu_busy_frac']:.2f}")
        rows = []
        for rt in sweep:
            rq = materialize(make_requests(sweep_requests, rt, seed + 7))
            mm = srv.run(rq)
            rows.append({"rate_per_s": rt, "requests": sweep_requests, **{
                k: mm[k] for k in ("p50_ms", "p99_ms", "targets_per_s", "mean_systems_per_round",
                                   "gpu_busy_frac", "makespan_s")},
                "errors": sum("error" in r.extra for r in rq)})
            if log:
                log(f"[serving] sweep {rt}/s: p50 {mm['p50_ms']:.1f} ms, p99 {mm['p99_ms']:.1f} ms,"
                    f" {mm['mean_systems_per_round']:.2f} systems/round, GPU busy "
                    f"{mm['gpu_busy_frac']:.2f} (rounds)")
    finally:
        srv.close()
    if baseline:
        # the sweep's arrivals served one request at a time as well
        for row in rows:
            rq = materialize(make_requests(sweep_requests, row["rate_per_s"], seed + 7))
            b = run_one_at_a_time(rq)
            row["one_at_a_time"] = {k: b[k] for k in ("p50_ms", "p99_ms", "targets_per_s",
                                                      "makespan_s")}
            if log:
                log(f"[serving] one at a time at {row['rate_per_s']}/s: p50 {b['p50_ms']:.1f} ms,"
                    f" p99 {b['p99_ms']:.1f} ms, {b['targets_per_s'] / 1e6:.1f}M targets/s")
    if rows:
        out["rate_sweep"] = rows
    if baseline:
        base = materialize(make_requests(requests, rate, seed))
        out["one_at_a_time"] = run_one_at_a_time(base)
        if log:
            b = out["one_at_a_time"]
            log(f"[serving] one at a time: p50 {b['p50_ms']:.1f} ms, p99 {b['p99_ms']:.1f} ms")
    return out


def main():
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="serving stream benchmark (one JSON line)")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=400.0)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--builders", type=int, default=4)
    ap.add_argument("--hbm-budget-gb", type=float, default=16.0)
    ap.add_argument("--no-baseline", action="store_true")
    ap.add_argument("--sweep", default="", help="comma-separated extra rates (req/s)")
    ap.add_argument("--sweep-requests", type=int, default=1000)
    ap.add_argument("--warmup-requests", type=int, default=64,
                    help="untimed stream (seed + 1) served first")
    a = ap.parse_args()
    sweep = [float(v) for v in a.sweep.split(",") if v]
    line = benchmark(a.requests, a.rate, a.seed, a.builders, int(a.hbm_budget_gb * (1 << 30)),
                     a.warmup_requests, not a.no_baseline, sweep, a.sweep_requests,
                     log=lambda m: print(m, file=sys.stderr, flush=True))
    line["workload"] = "serving stream"
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
