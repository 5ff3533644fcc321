"""Request loop for circuit transients on one GPU, with continuous batching.

A request is one circuit (fixed pattern) plus a number of Newton steps; every
step brings new values and a new right-hand side and wants x to rel err
<= 1e-6.  The loop:

  * prepares each request on arrival: host symbolic phase (COLAMD, pivots,
    pattern, levels, plan) in a CPU pool, and its device-memory estimate;
  * admits prepared requests as cohort slots and memory free up: the
    prepared requests waiting at that moment that share the oldest one's size
    class (n below 2k, below 8k, above) -- up to cohort_max of them, while
    their estimates fit the HBM budget beside the cohorts already running --
    are merged into one resident batch (Batch.adopt: their host structures
    move over, no symbolic phase is repeated);
  * runs every cohort on its own HIP stream: each step is one persistent
    refactor launch and one batched fp64 GMRES-IR over all its requests, and
    the cohorts in flight overlap on the device.  A request that has done all
    its steps is masked out of its cohort's later launches (set_active); the
    cohort's memory returns to the budget when its last request completes;
  * reports per-step latency (completion - submission; step 0 is submitted
    at arrival, step s when step s-1 completes) as p50 / p99, solves/s, the
    cohorts formed and the mean number of requests per launch.

Merging by size class keeps a 40k-unknown circuit from setting the step time
of 500-unknown ones (a step lasts as long as its slowest circuit), and
forming cohorts from requests that have not started yet means no batch is
ever rebuilt around requests in flight (a rebuild re-uploads every
structure in it).

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


SIZE_CLASSES = (2000, 8000)        # n below the first, below the second, the rest


def size_class(n: int) -> int:
    return sum(n >= b for b in SIZE_CLASSES)


class Server:
    """Cohort batching over concurrent resident batches (see the module text).

    slots        cohorts in flight at once (each on its own stream)
    cohort_max   most requests merged into one cohort
    gather_ms    how long a prepared request may wait for requests that
                 arrived before it was admitted and are still being prepared
    hbm_budget   bytes the running requests' device estimates may take together
    prec         factor storage of the steps ("fp32": the mixed path)
    """

    def __init__(self, hbm_budget: int, slots: int = 8, prec: str = "fp32", keep_results: bool = False,
                 prep_workers: int = 16, cohort_max: int = 32, gather_ms: float = 0.0, **opts):
        self.budget = int(hbm_budget)
        self.gather = gather_ms * 1e-3         # wait this long for requests still being prepared
        self.preparing = 0
        self.slots = slots
        self.cohort_max = cohort_max
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
        self.cohorts = 0
        self.launch_steps = 0
        self.circuit_steps = 0
        self.max_in_cohort = 0
        self.log_steps = False
        self.step_log = []                  # (cohort id, request ids, t_start, t_end) per step

    # -- host phase on arrival ---------------------------------------------
    def _prepare(self, r: Request):
        t_prep = time.perf_counter()
        try:
            c = r.circuit
            r.batch = Batch(Circuit(c.n, c.row, c.col, c.val, c.rhs, q=c.q), upload=False, **self.opts)
            r.estimate = r.batch.estimate_bytes()
            if r.estimate > self.budget:
                raise MemoryError(f"request {r.rid} needs {r.estimate} B > budget {self.budget} B")
        except Exception as e:          # reported, the loop goes on
            r.error = repr(e)
            if r.batch is not None:
                r.batch.close()
                r.batch = None
            with self.lock:
                self.preparing -= 1
            r.done.set()
            return
        with self.lock:
            r.ready_at = time.perf_counter()
            r.prep_s = r.ready_at - t_prep
            self.preparing -= 1
            self.ready.append(r)
            self.lock.notify_all()

    # -- one cohort on its own stream ------------------------------------------
    def _run_cohort(self, cid: int, group: list, held: int):
        B = None
        try:
            B = Batch.adopt([(r.batch, 0) for r in group])
            for r in group:
                r.batch = None                      # consumed by the cohort's batch
            vals = np.concatenate([r.circuit.val for r in group])
            rhs = np.concatenate([r.circuit.rhs for r in group])
            live = list(range(len(group)))
            while live:
                for i in live:
                    r = group[i]
                    v, b = r.step_inputs(r.step)
                    vals[B.aoff[i]:B.aoff[i + 1]] = v
                    rhs[B.voff[i]:B.voff[i + 1]] = b
                ts = time.perf_counter()
                B.set_active([r.step < r.steps for r in group])
                B.set_values(vals)
                rr = B.refactor(self.prec)
                x, sr = B.solve(rhs)
                if rr["rc"] != 0 or sr["rc"] != 0:
                    raise RuntimeError(f"cohort {cid} step failed: refactor {rr} solve {sr}")
                now = time.perf_counter()
                with self.lock:
                    self.launch_steps += 1
                    self.circuit_steps += len(live)
                    if self.log_steps:
                        self.step_log.append((cid, tuple(group[i].rid for i in live), ts, now))
                for i in live:
                    r = group[i]
                    r.lat.append(now - r.sub)
                    r.sub = now
                    r.reports.append((rr, sr))
                    if self.keep:
                        r.results.append(x[B.voff[i]:B.voff[i + 1]].copy())
                    r.step += 1
                    if r.step == r.steps:
                        r.done.set()
                live = [i for i in live if group[i].step < group[i].steps]
        except Exception as e:
            for r in group:
                if not r.done.is_set():
                    r.error = repr(e)
                    r.done.set()
        finally:
            for r in group:
                if r.batch is not None:
                    r.batch.close()
                    r.batch = None
            if B is not None:
                B.close()
            with self.lock:
                self.used -= held
                self.free_slots += 1
                self.lock.notify_all()

    def _take_cohort(self):
        """Under the lock: the next cohort (the oldest prepared request and the
        prepared requests of its size class behind it that fit), or None."""
        if not self.ready or self.free_slots == 0:
            return None
        if self.preparing > 0 and time.perf_counter() - self.ready[0].ready_at < self.gather:
            return None                     # requests still being prepared may join
        cls = size_class(self.ready[0].circuit.n)
        group, held, rest = [], 0, collections.deque()
        while self.ready:
            r = self.ready.popleft()
            if (len(group) < self.cohort_max and size_class(r.circuit.n) == cls
                    and self.used + held + r.estimate <= self.budget):
                group.append(r)
                held += r.estimate
            else:
                rest.append(r)
        self.ready = rest
        if not group:
            return None
        self.used += held
        self.peak = max(self.peak, self.used)
        self.free_slots -= 1
        return group, held

    def _admit_loop(self, total: int):
        admitted = 0
        while admitted < total:
            with self.lock:
                while True:
                    got = self._take_cohort()
                    if got is not None:
                        break
                    if self._stop:
                        return
                    self.lock.wait(0.005)
            group, held = got
            cid = self.cohorts
            self.cohorts += 1
            self.max_in_cohort = max(self.max_in_cohort, len(group))
            self.run_pool.submit(self._run_cohort, cid, group, held)
            admitted += len(group)

    def run(self, requests: list[Request]) -> dict:
        self._stop = False
        t0 = time.perf_counter()
        for r in requests:
            r.step = 0
            r.sub = t0 + r.arrival
        admit = threading.Thread(target=self._admit_loop, args=(len(requests),), daemon=True)
        admit.start()
        for r in sorted(requests, key=lambda q: q.arrival):
            wait = t0 + r.arrival - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            with self.lock:
                self.preparing += 1
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
        prep = np.array([getattr(r, "prep_s", 0.0) for r in requests]) * 1e3
        first = np.array([r.lat[0] for r in requests if r.lat]) * 1e3
        later = np.array([v for r in requests for v in r.lat[1:]]) * 1e3
        return {
            "requests": len(requests), "solves": solves, "wall_s": round(wall, 3),
            "solves_per_s": round(solves / wall, 2) if wall > 0 else 0.0,
            "p50_ms": round(float(np.percentile(lat, 50)), 3) if solves else None,
            "p99_ms": round(float(np.percentile(lat, 99)), 3) if solves else None,
            "max_rel_residual": float(max_res), "errors": len(errors),
            "peak_admitted_bytes": int(self.peak), "budget_bytes": self.budget,
            "cohorts": self.cohorts, "launch_steps": self.launch_steps,
            "mean_requests_per_launch": round(self.circuit_steps / max(1, self.launch_steps), 2),
            "max_requests_per_cohort": self.max_in_cohort,
            # where the tail comes from: the host symbolic phase on arrival
            # (step 0 includes it) against the steps of resident requests
            "prep_ms_p50": round(float(np.percentile(prep, 50)), 2) if prep.size else None,
            "prep_ms_max": round(float(prep.max()), 2) if prep.size else None,
            "first_step_p99_ms": round(float(np.percentile(first, 99)), 3) if first.size else None,
            "later_steps_p50_ms": round(float(np.percentile(later, 50)), 3) if later.size else None,
            "later_steps_p99_ms": round(float(np.percentile(later, 99)), 3) if later.size else None,
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
    ap.add_argument("--cohort-max", type=int, default=32)
    ap.add_argument("--gather-ms", type=float, default=1.0)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--seed", type=int, default=4)
    args = ap.parse_args(argv)
    from . import device_sync

    for rate in args.rate:
        reqs = make_requests(args.requests, rate, nmax=args.nmax, seed=args.seed)
        srv = Server(int(args.budget_gb * 2**30), slots=args.slots, cohort_max=args.cohort_max,
                     gather_ms=args.gather_ms, keep_results=True)
        st = srv.run(reqs)
        srv.close()
        device_sync()
        st["rate"] = rate
        st["rel_err_max_sample"] = check_against_cpu(reqs)
        print(json.dumps(st), flush=True)


if __name__ == "__main__":
    main()
