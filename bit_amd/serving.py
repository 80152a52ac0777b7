"""Continuous-batching request server for independent treecode requests.

Requests (potential evaluations of independent particle systems, and sphere
boundary-integral GMRES solves) arrive over time.  The server:

* admits arrived requests in order while their estimated device memory fits
  the HBM budget;
* builds each admitted request's tree, moments and interaction lists on a pool
  of builder threads (each with its own context/stream, a few host threads for
  the tree);
* runs a GPU loop that, each round, takes every ready evaluation and every
  pending GMRES matrix-vector product and evaluates all of them in one mixed
  launch per parameter group (bit_eval_potential_multi) -- requests arriving
  while a round runs join the next one, and a GMRES solve occupies a slot for
  as many rounds as it needs products;
* keeps each solve's tree and lists resident between its products (charges are
  swapped in place, bit_system_set_charges).

Every target's sums are formed the same way alone and in a mix, so results do
not depend on what else was on the GPU (tests/test_gpu_serving.py).
"""
from __future__ import annotations

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
    result: np.ndarray | None = None
    iters: int = 0
    relres: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def est_bytes(self) -> int:
        return BYTES_PER_PARTICLE * self.n + BYTES_FIXED

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


def _particles(r: Request):
    if r.kind == "eval":
        P = workloads.make(r.layout, r.n, r.seed)
        return P.x, P.y, P.z, P.q, None
    V, F = bie.icosphere(r.level)
    pts, area, h = bie.collocation(V, F)
    r.delta = h
    w = area / (4.0 * np.pi)
    r.extra["f"] = bie.random_rhs(8, r.seed % 100000)(pts)
    r.extra["self"] = float(np.exp(-r.kappa * h) / h)
    return pts[:, 0].copy(), pts[:, 1].copy(), pts[:, 2].copy(), w, w


def solve_alone(ctx, r: Request, precision="fp32"):
    """The request evaluated by itself (reference result for the mix tests)."""
    x, y, z, q, w = _particles(r)
    s = bit_amd.System(ctx, x, y, z, q, r.maxmemb)
    try:
        args = (r.nterms, r.acc, r.kappa, r.delta ** 2, None, precision)
        if r.kind == "eval":
            return s.potential(*args)
        f, selfc = r.extra["f"], r.extra["self"]

        def matvec(v):  # the server's operator and GMRES core, one system alone
            s.set_charges(w * v)
            return (0.5 + selfc * w) * v + s.potential(*args)

        sig, it, rr = bit_amd.gmres(s.n, matvec, f, None, 50, r.tol, 500)
        return sig
    finally:
        s.close()


class _Job:
    __slots__ = ("req", "system", "params", "out", "done")

    def __init__(self, req, system, params):
        self.req, self.system, self.params = req, system, params
        self.out = None
        self.done = threading.Event()


class Server:
    def __init__(self, device: int = 0, hbm_budget: int = 16 << 30, builders: int = 8,
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
        work: "queue.Queue[_Job]" = queue.Queue()
        inflight = [0]
        cv = threading.Condition()
        pending = list(reversed(reqs))
        finished = [0]
        total = len(reqs)
        t0 = time.perf_counter()
        pool = ThreadPoolExecutor(max_workers=self.builders)
        solvers = []
        rounds = []

        def now():
            return time.perf_counter() - t0

        def release(r, s):
            s.close()
            with cv:
                inflight[0] -= r.est_bytes
                finished[0] += 1
                cv.notify_all()

        def fail(r, e):
            r.extra["error"] = repr(e)
            r.t_done = now()
            with cv:
                inflight[0] -= r.est_bytes
                finished[0] += 1
                cv.notify_all()

        def build(r):
            try:
                ctx = self._builder_ctx()
                x, y, z, q, w = _particles(r)
                s = bit_amd.System(ctx, x, y, z, q, r.maxmemb)
                prm = r.params(self.precision)
                s.prepare(r.nterms, r.acc, r.kappa, r.delta ** 2, None, self.precision)
            except Exception as e:  # a failed request must not stall the loop
                fail(r, e)
                return
            r.t_ready = now()
            if r.kind == "eval":
                work.put(_Job(r, s, prm))
                return
            th = threading.Thread(target=gmres_request, args=(r, s, w, prm), daemon=True)
            solvers.append(th)
            th.start()

        def gmres_request(r, s, w, prm):
            f, selfc = r.extra["f"], r.extra["self"]

            def matvec(v):
                s.set_charges(w * v)
                job = _Job(r, s, prm)
                work.put(job)
                job.done.wait()
                return (0.5 + selfc * w) * v + job.out

            try:
                sig, it, rr = bit_amd.gmres(s.n, matvec, f, None, 50, r.tol, 500)
            except Exception as e:
                s.close()
                fail(r, e)
                return
            r.result, r.iters, r.relres = sig, it, rr
            r.t_done = now()
            release(r, s)

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
        while finished[0] < total:
            try:
                first = work.get(timeout=0.005)
            except queue.Empty:
                continue
            jobs = [first]
            while len(jobs) < self.max_round:
                try:
                    jobs.append(work.get_nowait())
                except queue.Empty:
                    break
            outs, ms = bit_amd.eval_multi(self.gpu_ctx, [j.system for j in jobs],
                                          [j.params for j in jobs])
            rounds.append((len(jobs), sum(j.system.n for j in jobs), ms))
            t = now()
            for j, o in zip(jobs, outs):
                if j.req.kind == "eval":
                    j.req.result = o
                    j.req.t_done = t
                    release(j.req, j.system)
                else:
                    j.out = o
                    j.done.set()
        adm.join()
        pool.shutdown(wait=True)
        for th in solvers:
            th.join()
        makespan = max(r.t_done for r in reqs) - reqs[0].arrival
        return summarize(reqs, makespan, rounds)


def run_one_at_a_time(requests, device: int = 0, precision: str = "fp32",
                      realtime: bool = True):
    """Baseline: the same arrivals served one request at a time (one system,
    one evaluation or solve, on one stream)."""
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


def main():
    import argparse
    import json
    ap = argparse.ArgumentParser(description="serving stream benchmark (one JSON line)")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=400.0)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--builders", type=int, default=8)
    ap.add_argument("--hbm-budget-gb", type=float, default=16.0)
    ap.add_argument("--no-baseline", action="store_true")
    ap.add_argument("--warmup-requests", type=int, default=64,
                    help="untimed stream (seed + 1) served first")
    a = ap.parse_args()
    reqs = make_requests(a.requests, a.rate, a.seed)
    srv = Server(hbm_budget=int(a.hbm_budget_gb * (1 << 30)), builders=a.builders)
    if a.warmup_requests > 0:
        srv.run(make_requests(a.warmup_requests, a.rate, a.seed + 1))
    m = srv.run(reqs)
    srv.close()
    errors = sum("error" in r.extra for r in reqs)
    line = {"workload": "serving stream", "requests": a.requests, "rate_per_s": a.rate,
            "seed": a.seed, "warmup_requests": a.warmup_requests, "server": m,
            "errors": errors}
    if not a.no_baseline:
        base = make_requests(a.requests, a.rate, a.seed)
        line["one_at_a_time"] = run_one_at_a_time(base)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
