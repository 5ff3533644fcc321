"""Request loop for circuit transients on one GPU.

A request is one circuit (fixed pattern) plus a number of Newton steps; every
step brings new values and a new right-hand side and wants x to rel err
<= 1e-6.  The loop:

  * prepares each request on arrival: host symbolic phase (COLAMD, pivots,
    pattern, levels, plan) in a CPU pool;
  * admits prepared requests in arrival order when a slot is free and the
    request's estimated device memory fits the free-HBM budget;
  * keeps an admitted request's structure, factors and Krylov workspace
    resident for all its steps, on its own HIP stream, so admitted requests
    overlap on the device; the memory returns to the budget when the
    request completes;
  * reports per-step latency (completion - submission; step 0 is submitted
    at arrival, step s when step s-1 completes) as p50 / p99, and solves/s.

    python -m luspice_amd.serve --rate 1000 --requests 200
"""
from __future__ import annotations

import argparse
import collections
import json
import math
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import Batch, Circuit, colamd, cpu_lu_ref, mna, perturb


class Request:
    def __init__(self, rid: int, arrival: float, circuit: Circuit, steps: int, seed: int):
        self.rid = rid
        self.arrival = arrival
        self.circuit = circuit
        self.steps = steps
        self.seed = seed
        self.batch = None
        self.estimate = 0
        self.lat = []            # per-step latency, seconds
        self.results = []        # x per step (kept when the server keeps results)
        self.reports = []
        self.error = None
        self.done = threading.Event()

    def step_inputs(self, s: int):
        return perturb(self.seed, s, self.circuit.val, self.circuit.n, 0.1)


def make_requests(count: int, rate: float, nmin: int = 200, nmax: int = 40000, smin: int = 5,
                  smax: int = 50, seed: int = 4) -> list[Request]:
    """Poisson arrivals at ``rate`` requests/s, n log-uniform, steps uniform."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, count)
    arrivals = np.cumsum(gaps) - gaps[0]
    ns = np.rint(np.exp(rng.uniform(math.log(nmin), math.log(nmax), count))).astype(int)
    steps = rng.integers(smin, smax + 1, count)
    out = []
    for i in range(count):
        c = mna("sized", n=int(ns[i]), lr=0.02, seed=seed * 1_000_003 + i)
        c.q = colamd(c)
        out.append(Request(i, float(arrivals[i]), c, int(steps[i]), seed * 7_919 + i))
    return out


class Server:
    def __init__(self, hbm_budget: int, slots: int = 8, prec: str = "fp32", keep_results: bool = False,
                 prep_workers: int = 8, **opts):
        self.budget = int(hbm_budget)
        self.slots = slots
        self.prec = prec
        self.keep = keep_results
        self.opts = opts
        self.used = 0
        self.peak = 0
        self.lock = threading.Condition()
        self.ready = collections.deque()
        self.free_slots = slots
        self.prep_pool = ThreadPoolExecutor(prep_workers)
        self.run_pool = ThreadPoolExecutor(slots)

    # -- request life cycle --------------------------------------------------
    def _prepare(self, r: Request):
        try:
            c = r.circuit
            r.batch = Batch(Circuit(c.n, c.row, c.col, c.val, c.rhs, q=c.q), upload=False, **self.opts)
            r.estimate = r.batch.estimate_bytes()
            if r.estimate > self.budget:
                raise MemoryError(f"request {r.rid} needs {r.estimate} B > budget {self.budget} B")
        except Exception as e:          # reported, the loop goes on
            r.error = repr(e)
            r.done.set()
            return
        with self.lock:
            self.ready.append(r)
            self.lock.notify_all()

    def _run(self, r: Request, t0: float):
        try:
            B = r.batch
            B.upload()
            sub = t0 + r.arrival
            for s in range(r.steps):
                vals, rhs = r.step_inputs(s)
                B.set_values(vals)
                rr = B.refactor(self.prec)
                x, sr = B.solve(rhs)
                now = time.perf_counter()
                r.lat.append(now - sub)
                sub = now
                r.reports.append((rr, sr))
                if self.keep:
                    r.results.append(x)
        except Exception as e:
            r.error = repr(e)
        finally:
            if r.batch is not None:
                r.batch.close()
                r.batch = None
            with self.lock:
                self.used -= r.estimate
                self.free_slots += 1
                self.lock.notify_all()
            r.done.set()

    def _admit_loop(self, total: int, t0: float):
        admitted = 0
        while admitted < total:
            with self.lock:
                while True:
                    if self.ready and self.free_slots > 0 and self.used + self.ready[0].estimate <= self.budget:
                        r = self.ready.popleft()
                        self.used += r.estimate
                        self.peak = max(self.peak, self.used)
                        self.free_slots -= 1
                        break
                    if self._stop and not self.ready:
                        return
                    self.lock.wait(0.05)
            self.run_pool.submit(self._run, r, t0)
            admitted += 1

    def run(self, requests: list[Request]) -> dict:
        self._stop = False
        t0 = time.perf_counter()
        admit = threading.Thread(target=self._admit_loop, args=(len(requests), t0), daemon=True)
        admit.start()
        for r in sorted(requests, key=lambda q: q.arrival):
            wait = t0 + r.arrival - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            self.prep_pool.submit(self._prepare, r)
        for r in requests:
            r.done.wait()
        with self.lock:
            self._stop = True
            self.lock.notify_all()
        admit.join(timeout=5)
        wall = time.perf_counter() - t0
        lat = np.array([v for r in requests for v in r.lat]) * 1e3
        solves = int(lat.size)
        errors = [r.error for r in requests if r.error]
        max_res = max((sr["max_rel_res"] for r in requests for _, sr in r.reports), default=0.0)
        return {
            "requests": len(requests), "solves": solves, "wall_s": round(wall, 3),
            "solves_per_s": round(solves / wall, 2) if wall > 0 else 0.0,
            "p50_ms": round(float(np.percentile(lat, 50)), 3) if solves else None,
            "p99_ms": round(float(np.percentile(lat, 99)), 3) if solves else None,
            "max_rel_residual": float(max_res), "errors": len(errors),
            "peak_admitted_bytes": int(self.peak), "budget_bytes": self.budget,
        }

    def close(self):
        self.prep_pool.shutdown(wait=True)
        self.run_pool.shutdown(wait=True)


def check_against_cpu(requests, count: int = 16) -> float:
    """Max rel error of the last kept step of ``count`` requests vs the fp64
    CPU LU (cpu_lu_ref: threshold partial pivoting, same column order)."""
    worst = 0.0
    for r in requests[:count]:
        if not r.results:
            continue
        s = len(r.results) - 1
        vals, rhs = r.step_inputs(s)
        c = r.circuit
        rc, xr, _, _ = cpu_lu_ref(Circuit(c.n, c.row, c.col, vals, rhs), c.q)
        worst = max(worst, float(np.linalg.norm(r.results[s] - xr) / np.linalg.norm(xr)))
    return worst


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, nargs="+", default=[100, 1000, 10000])
    ap.add_argument("--requests", type=int, default=120)
    ap.add_argument("--nmax", type=int, default=40000)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--seed", type=int, default=4)
    args = ap.parse_args(argv)
    from . import device_sync

    for rate in args.rate:
        reqs = make_requests(args.requests, rate, nmax=args.nmax, seed=args.seed)
        srv = Server(int(args.budget_gb * 2**30), slots=args.slots, keep_results=True)
        st = srv.run(reqs)
        srv.close()
        device_sync()
        st["rate"] = rate
        st["rel_err_max_sample"] = check_against_cpu(reqs)
        print(json.dumps(st), flush=True)


if __name__ == "__main__":
    main()
