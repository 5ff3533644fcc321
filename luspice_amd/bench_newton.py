"""Batched Newton re-factorisation benchmark (the circuit mix of many uneven n).

    python -m luspice_amd.bench_newton [--circuits 512] [--steps 20] [--nmin 100] [--nmax 5000]

Workload: ``circuits`` MNA circuits, n log-uniform in [nmin, nmax] (seed 3),
fixed pattern each; every Newton step scales the values by lognormal(0, 0.1)
and draws b ~ U[-1, 1].  One unit = one circuit refactored and solved to
rel err <= 1e-6.

  batched: all circuits in one resident batch, fp32 factors, one persistent
           refactor launch + one batched fp64 GMRES-IR per step, inputs in HBM,
           the whole step replayed as one HIP graph (--no-graph: eager)
  plain:   one batch per circuit, fp64 factors, host inputs each step (timed
           over --plain-steps steps)

Accuracy: every circuit's solution at the last batched step is compared with
the fp64 CPU LU of that step's matrix (cpu_lu_ref: partial pivoting, same
column order), as relative L2 error.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--circuits", type=int, default=512)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--nmin", type=int, default=100)
    ap.add_argument("--nmax", type=int, default=5000)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--check", type=int, default=-1, help="circuits checked vs the CPU LU (-1 = all)")
    ap.add_argument("--no-graph", action="store_true", help="eager launches instead of the HIP graph")
    ap.add_argument("--prec", default="fp32", choices=["fp32", "fp64"], help="factor storage of the batched path")
    args = ap.parse_args(argv)

    from . import device_sync
    from . import workloads as W

    def log(m):
        print(m, file=sys.stderr, flush=True)

    t0 = time.time()
    circ = W.newton_mix(args.circuits, args.nmin, args.nmax, args.seed)
    ntot = sum(c.n for c in circ)
    log(f"[newton] {len(circ)} circuits, total n {ntot}, generated+colamd {time.time() - t0:.1f}s")

    t0 = time.time()
    NB = W.NewtonBatched(circ, args.steps + args.warmup, args.seed)
    tails = [NB.B.info(i)["tail_T"] for i in range(len(circ))]
    log(f"[newton] batch built {time.time() - t0:.1f}s, dense tails {sum(t > 0 for t in tails)}, max T {max(tails)}")
    for s in range(args.warmup):
        NB.run_step(s, prec=args.prec, graph=not args.no_graph)
    device_sync()
    lat = []
    reps = []
    t1 = time.perf_counter()
    for s in range(args.warmup, args.warmup + args.steps):
        ts = time.perf_counter()
        rr, sr = NB.run_step(s, prec=args.prec, graph=not args.no_graph)
        lat.append((time.perf_counter() - ts) * 1e3)
        reps.append((rr, sr))
    device_sync()
    wall_b = time.perf_counter() - t1
    x_last = NB.solution()
    units_b = len(circ) * args.steps / wall_b
    log(f"[newton] batched: {units_b:.1f} units/s, step p50 {np.percentile(lat, 50):.2f} ms")

    # accuracy of the last batched step against the fp64 CPU LU (the
    # library's cpu_lu_ref: threshold partial pivoting, fp64)
    from . import Circuit, cpu_lu_ref
    last = args.warmup + args.steps - 1
    vals, rhs = NB.host[last]
    aoff = np.cumsum([0] + [c.nnz for c in circ])
    voff = np.cumsum([0] + [c.n for c in circ])
    idx = range(len(circ)) if args.check < 0 else np.linspace(0, len(circ) - 1, args.check).astype(int)
    errs = []
    for i in idx:
        c = circ[i]
        ci = Circuit(c.n, c.row, c.col, vals[aoff[i]:aoff[i + 1]], rhs[voff[i]:voff[i + 1]])
        rc, xr, _, _ = cpu_lu_ref(ci, c.q)
        xi = x_last[voff[i]:voff[i + 1]]
        errs.append(float(np.linalg.norm(xi - xr) / np.linalg.norm(xr)))
    errs = np.array(errs)
    log(f"[newton] rel err vs fp64 CPU LU: max {errs.max():.3e} median {np.median(errs):.3e}")

    # plain fp64 path, one matrix per launch
    t0 = time.time()
    NP = W.NewtonPlain(circ, args.plain_steps + 1, host_inputs=NB.host[:args.plain_steps + 1])
    log(f"[newton] plain batches built {time.time() - t0:.1f}s")
    NP.run_step(0)
    device_sync()
    t1 = time.perf_counter()
    for s in range(1, args.plain_steps + 1):
        xp, preps = NP.run_step(s)
    device_sync()
    wall_p = time.perf_counter() - t1
    units_p = len(circ) * args.plain_steps / wall_p
    log(f"[newton] plain fp64: {units_p:.1f} units/s")

    inner = [r[1]["inner_total"] for r in reps]
    fallback = sum(r[1]["fallback"] for r in reps)
    line = {
        "metric": ("batched Newton refactor+solve units/s at rel err <=1e-6 ("
                   + ("mixed" if args.prec == "fp32" else "fp64") + " batched vs plain fp64)"),
        "value": round(units_b, 2),
        "unit": "circuit refactor+solve/s",
        "plain_fp64": round(units_p, 2),
        "batched_over_plain": round(units_b / units_p, 2),
        "circuits": len(circ), "total_n": ntot, "steps": args.steps,
        "step_p50_ms": round(float(np.percentile(lat, 50)), 3),
        "step_p99_ms": round(float(np.percentile(lat, 99)), 3),
        "rel_err_max": float(errs.max()), "rel_err_median": float(np.median(errs)),
        "pass_1e-6": bool(errs.max() <= 1e-6 and all(r[1]["rc"] == 0 for r in reps)),
        "ir_inner_median": int(np.median(inner)), "fallback_steps": int(fallback),
        "dense_tails": int(sum(t > 0 for t in tails)), "max_tail_T": int(max(tails)),
        "hip_graph": not args.no_graph,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
