"""Batched Newton re-factorisation benchmark (the circuit mix of many uneven n).

    python -m luspice_amd.bench_newton [--circuits 512] [--steps 20] [--nmin 100] [--nmax 5000]
                                       [--precs fp32 fp64]

Workload: ``circuits`` MNA circuits, n log-uniform in [nmin, nmax] (seed 3),
fixed pattern each; every Newton step scales the values by lognormal(0, 0.1)
and draws b ~ U[-1, 1].  One unit = one circuit refactored and solved to
rel err <= 1e-6.

  batched: all circuits in one resident batch, one persistent refactor launch
           + one batched fp64 GMRES-IR per step, inputs in HBM, the whole step
           replayed as one HIP graph (--no-graph: eager); factor storage fp32
           (the mixed path) and fp64, each on a batch of its own; with
           --ir-methods 1 the solve is fused Richardson steps instead (one
           launch per step for the whole batch, GMRES-IR for what is left)
  plain:   one batch per circuit, fp64 factors, host inputs each step (timed
           over --plain-steps steps): the one-matrix-per-launch fp64 path

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


def _log(m):
    print(m, file=sys.stderr, flush=True)


def _errors(circ, NB, last, check):
    """Relative L2 error of each checked circuit's solution at step `last`
    against the fp64 CPU LU of that step's matrix."""
    from . import Circuit, cpu_lu_ref
    x_last = NB.solution()
    vals, rhs = NB.host[last]
    aoff = np.cumsum([0] + [c.nnz for c in circ])
    voff = np.cumsum([0] + [c.n for c in circ])
    idx = range(len(circ)) if check < 0 else np.linspace(0, len(circ) - 1, check).astype(int)
    errs = []
    for i in idx:
        c = circ[i]
        ci = Circuit(c.n, c.row, c.col, vals[aoff[i]:aoff[i + 1]], rhs[voff[i]:voff[i + 1]])
        rc, xr, _, _ = cpu_lu_ref(ci, c.q)
        xi = x_last[voff[i]:voff[i + 1]]
        errs.append(float(np.linalg.norm(xi - xr) / np.linalg.norm(xr)))
    return np.array(errs)


def batched(circ, prec, steps, warmup, seed, graph=True, check=-1, host=None, ir_method=0):
    """Runs `steps` timed Newton steps (after `warmup`) of the whole mix in one
    resident batch with `prec` factor storage.  Returns the result record and
    the step inputs (reused by the plain path)."""
    from . import device_sync
    from . import workloads as W
    t0 = time.time()
    NB = W.NewtonBatched(circ, steps + warmup, seed, host_inputs=host, ir_method=ir_method)
    tails = [NB.B.info(i)["tail_T"] for i in range(len(circ))]
    _log(f"[newton] {prec} batch built {time.time() - t0:.1f}s, dense tails "
         f"{sum(t > 0 for t in tails)}, max T {max(tails)}")
    repiv = 0
    for s in range(warmup):
        rr, _ = NB.run_step(s, prec=prec, graph=graph)
        repiv += rr.get("repivoted", 0)
    device_sync()
    lat, reps = [], []
    t1 = time.perf_counter()
    for s in range(warmup, warmup + steps):
        ts = time.perf_counter()
        rr, sr = NB.run_step(s, prec=prec, graph=graph)
        lat.append((time.perf_counter() - ts) * 1e3)
        reps.append((rr, sr))
    device_sync()
    wall = time.perf_counter() - t1
    units = len(circ) * steps / wall
    _log(f"[newton] batched {prec}: {units:.1f} units/s, step p50 {np.percentile(lat, 50):.2f} ms")
    errs = _errors(circ, NB, warmup + steps - 1, check)
    if errs.size:
        _log(f"[newton] {prec} rel err vs fp64 CPU LU: max {errs.max():.3e} median {np.median(errs):.3e}")
    else:                                      # --check 0: throughput only, nothing compared
        errs = np.full(1, np.nan)
    inner = [r[1]["inner_total"] for r in reps]
    rec = {
        "value": round(units, 2), "unit": "circuit refactor+solve/s", "prec": prec,
        "ms_per_step": round(1e3 * wall / steps, 3),
        "step_p50_ms": round(float(np.percentile(lat, 50)), 3),
        "step_p99_ms": round(float(np.percentile(lat, 99)), 3),
        "rel_err_max": float(errs.max()), "rel_err_median": float(np.median(errs)),
        "pass_1e-6": bool(errs.max() <= 1e-6 and all(r[1]["rc"] == 0 for r in reps)),   # False unchecked
        "ir_inner_median": int(np.median(inner)),
        "fallback_steps": int(sum(r[1]["fallback"] > 0 for r in reps)),
        "repivoted": int(repiv + sum(r[0].get("repivoted", 0) for r in reps)),
        "bad_pivots": int(sum(r[0].get("bad_pivots", 0) for r in reps)),
        "dense_tails": int(sum(t > 0 for t in tails)), "max_tail_T": int(max(tails)),
        "hip_graph": bool(graph),
        "ir": "fused Richardson steps" if ir_method == 1 else "GMRES-IR",
    }
    host = NB.host
    NB.close()
    return rec, host


def plain(circ, host, plain_steps):
    """The one-matrix-per-launch fp64 path on the first plain_steps + 1 step
    inputs (the first step untimed)."""
    from . import device_sync
    from . import workloads as W
    t0 = time.time()
    NP = W.NewtonPlain(circ, plain_steps + 1, host_inputs=host[:plain_steps + 1])
    _log(f"[newton] plain batches built {time.time() - t0:.1f}s")
    NP.run_step(0)
    device_sync()
    t1 = time.perf_counter()
    ok = True
    for s in range(1, plain_steps + 1):
        _, preps = NP.run_step(s)
        ok = ok and all(r[1]["rc"] == 0 for r in preps)
    device_sync()
    wall = time.perf_counter() - t1
    NP.close()
    units = len(circ) * plain_steps / wall
    _log(f"[newton] plain fp64: {units:.1f} units/s")
    return {"value": round(units, 2), "unit": "circuit refactor+solve/s", "steps": plain_steps,
            "all_converged": bool(ok)}


def compare(ncirc=512, steps=20, warmup=1, plain_steps=2, nmin=100, nmax=5000, seed=3,
            precs=("fp32", "fp64"), graph=True, check=-1, ir_methods=(0,)):
    """Batched (each precision in `precs`) against plain fp64 on one mix."""
    from . import workloads as W
    t0 = time.time()
    circ = W.newton_mix(ncirc, nmin, nmax, seed)
    ntot = sum(c.n for c in circ)
    _log(f"[newton] {len(circ)} circuits, total n {ntot}, generated+colamd {time.time() - t0:.1f}s")
    out = {"workload": f"{len(circ)} MNA circuits, n log-uniform in [{nmin}, {nmax}], seed {seed}, "
                       f"{steps} Newton steps (values x lognormal(0, 0.1), b ~ U[-1, 1])",
           "circuits": len(circ), "total_n": ntot, "steps": steps}
    host = None
    keys = []
    for irm in ir_methods:
        for prec in precs:
            rec, host = batched(circ, prec, steps, warmup, seed, graph=graph, check=check, host=host,
                                ir_method=irm)
            key = ("mixed" if prec == "fp32" else prec) + ("_fused" if irm == 1 else "")
            out[key] = rec
            keys.append(key)
    p = plain(circ, host, plain_steps) if plain_steps > 0 else None
    out["plain_fp64"] = p
    for key in keys:
        if p:
            out[key]["over_plain"] = round(out[key]["value"] / p["value"], 2)
    if "mixed" in out and "fp64" in out:
        out["mixed_over_fp64"] = round(out["mixed"]["value"] / out["fp64"]["value"], 3)
    return out


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
    ap.add_argument("--precs", nargs="+", default=["fp32", "fp64"], choices=["fp32", "fp64"],
                    help="factor storage of the batched runs")
    ap.add_argument("--ir-methods", nargs="+", type=int, default=[0], choices=[0, 1],
                    help="0: GMRES-IR; 1: fused Richardson steps (one launch per step for the batch)")
    ap.add_argument("--calib-mb", type=int, default=0,
                    help="size of a device copy issued first (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args(argv)
    if args.calib_mb > 0:
        from . import DeviceArray, device_copy, device_sync
        nb = args.calib_mb << 20
        ca, cb = DeviceArray(np.zeros(nb // 8)), DeviceArray(np.zeros(nb // 8))
        device_copy(cb.ptr, ca.ptr, nb)
        device_sync()
        ca.close()
        cb.close()
        _log(f"[newton] calibration copy {nb} B")
    out = compare(args.circuits, args.steps, args.warmup, args.plain_steps, args.nmin, args.nmax,
                  args.seed, tuple(args.precs), not args.no_graph, args.check, tuple(args.ir_methods))
    out["metric"] = "batched Newton refactor+solve units/s at rel err <=1e-6 (batched vs plain fp64)"
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
