"""Headline benchmark: refactor+solve/sec at rel err <= 1e-6 on one MI355X.

Workload (the single-GPU configuration the metric is quoted on): the
316x316 power-grid MNA matrix, n = 99906, seed 1.  One step = one numeric
refactorization of the resident matrix (symbolic structure reused, as in a
Newton loop) + one fp64 GMRES-IR solve to 1e-10 true residual, with the
inputs (values, RHS) already in HBM.

  mixed path (value): fp32 factor storage, fp64 pivot test / work vectors,
                      fp64 GMRES-IR (SELL-C-64 residual, level-scheduled
                      trisolve preconditioner)
  fp64 path:          fp64 factor storage, same kernels, one matrix per launch

Every result is checked against the fp64 CPU solution of the reference's own
splu (oracle/_ref, compiled from the reference sources) or, when that build
is absent, the oracle's C restatement; the line reports the relative error.

Beside the headline the line carries:
  cpu_baseline            the CPU fp64 path (splu factor + solve), 1 thread
  cpu_baseline_all_cores  the same on every CPU of this process's share at
                          once (independent solves), core count and model
  cpu_mixed_c0            the reference's CPU mixed path (fp32 factors +
                          fp64 GMRES-IR) on the 1020-unknown circuit
  batched_newton          512 circuits (n in [100, 5000], seed 3) x 20 Newton
                          steps: the resident batched path with fp32 and
                          with fp64 factors against the plain
                          one-matrix-per-launch fp64 path, rel err of every
                          circuit vs the fp64 CPU LU, step p50/p99

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--newton-circuits 512] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, published


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter runs
    (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this command, calibrated by
    the known-size copy), or None when no such measurement is committed."""
    path = os.path.join(ROOT, "profiles", "traffic_c1.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get(kernel, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def tri_roofline(tri_b, tri_ms):
    """HBM roofline of one trisolve pair (forward and backward launches with
    their dense-block solves) on the fp32 factors: algorithmic bytes over the
    pair's event time; traffic from the counter runs of both sparse launches."""
    ach = tri_b / (tri_ms * 1e-3) / 1e9 if tri_ms > 0 else 0.0
    tf = measured_traffic("lsp::trisolve_fwd_kernel<float>")
    tb = measured_traffic("lsp::trisolve_bwd_kernel<float>")
    return {"kernel": "trisolve pair (trisolve_fwd/bwd_kernel<float> + tail_trsv)", "bound": "hbm",
            "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 5),
            "traffic": (tf + tb) if (tf is not None and tb is not None) else None,
            "traffic_covers": "the two sparse launches (dense-block solves not counted)",
            "algorithmic_bytes": int(tri_b), "duration_ms": round(tri_ms, 4)}


def algorithmic_bytes(inf, nnz, n, vF):
    """Per-launch algorithmic bytes (compulsory traffic) of each kernel."""
    nLU = inf["lnnz"] + inf["unnz"]
    fac = nnz * (8 + 4) + 4 * (n + 1) + nLU * 4 + 8 * (n + 1) + nLU * vF
    tri = nLU * (vF + 4) + 8 * (n + 1) + 8 * n + 8 * n + 8 * n
    spmv = nnz * (8 + 4) + 4 * (n + 1) + 8 * n + 8 * n
    return fac, tri, spmv


def host_cpus():
    """CPUs this process may run on (its affinity mask), at most 16: the GPU
    box gives one GPU's process a 16-CPU share whatever the machine holds."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_all_cores(O, c, q, which, workers, runs_per_worker=2):
    """The CPU fp64 path on every worker thread at once: each thread factors
    and solves its own copy of the matrix (independent solves, as a CPU
    Newton farm would); the oracle's C releases the GIL, so the threads run
    in parallel.  Returns (solves/s over the wall time, wall seconds)."""
    from concurrent.futures import ThreadPoolExecutor

    def work(_):
        for _r in range(runs_per_worker):
            rc, _x, _ln, _un = O.cpu_lu_ref(c.n, c.row, c.col, c.val, q, c.rhs, which=which)
            assert rc == -1, rc
        return runs_per_worker

    t1 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        done = sum(ex.map(work, range(workers)))
    wall = time.perf_counter() - t1
    return done / wall, wall


def cpu_mixed_c0(L, O, repeats=20):
    """The reference's CPU mixed path on the 1020-unknown circuit
    (bench/test_lu_gpu_f32.cu:540-715): fp32-storage splu factors (the
    reference's own splu compiled with fp32 values when oracle/_ref is built,
    else the oracle's fp64 factors rounded to fp32) preconditioning fp64
    GMRES-IR to the 1e-10 true residual; timed factor + solve, 1 thread."""
    c = L.mna("c0", seed=0)
    q = L.colamd(c)
    which = "ref32" if O.ref_available() else "oracle"
    rc, xr, _, _ = O.cpu_lu_ref(c.n, c.row, c.col, c.val, q, c.rhs,
                                which="ref64" if O.ref_available() else "oracle")
    times, err, info = [], None, None
    for _ in range(repeats):
        t1 = time.perf_counter()
        F = O.lu_factor(c.n, c.row, c.col, c.val, q, which=which)
        F32 = dict(F, Lx=F["Lx"].astype(np.float32), Ux=F["Ux"].astype(np.float32))
        x, info = O.gmres_ir(c.n, c.row, c.col, c.val, c.rhs, F32, q)
        times.append(time.perf_counter() - t1)
        err = float(np.linalg.norm(x - xr) / np.linalg.norm(xr))
    return {"value": round(1.0 / float(np.median(times)), 2), "unit": "refactor+solve/s", "cores": 1,
            "kind": "reference" if which == "ref32" else "port",
            "sample": f"{repeats} runs on c0 (n={c.n}, seed 0): {which} fp32-storage factor + the "
                      f"oracle's GMRES-IR restatement (fp64 residual), median",
            "rel_err": err, "ir_inner": int(info["inner"]), "ir_res": float(info["res"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=5, help="CPU baseline factor+solve runs")
    ap.add_argument("--calib-mb", type=int, default=64,
                    help="size of the device copy issued before warm-up (counter calibration)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-workers", type=int, default=0,
                    help="threads of the all-cores CPU leg (0 = the CPUs this process may use, at most 16)")
    ap.add_argument("--newton-circuits", type=int, default=512,
                    help="circuits of the batched Newton comparison (0 = skip it)")
    ap.add_argument("--newton-steps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the path runs on one GPU's share); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import luspice_amd as L
    from oracle import oracle as O

    t0 = time.time()
    c = L.mna("c1", seed=1)
    q = L.colamd(c)
    log(f"[bench] c1 n={c.n} nnz={c.nnz} generated+colamd in {time.time() - t0:.1f}s")
    circ = L.Circuit(c.n, c.row, c.col, c.val, c.rhs, q=q)

    t0 = time.time()
    B = L.Batch(circ)
    inf = B.info(0)
    log(f"[bench] symbolic+plan+upload {time.time() - t0:.1f}s  lnnz={inf['lnnz']} "
        f"unnz={inf['unnz']} updates={inf['n_updates']} levels={inf['nlevels']}")
    # the right-hand side lives in the batch's resident arena for the whole run
    B.set_rhs(c.rhs)
    L.device_sync()
    # one plain device-to-device copy of known size before anything is timed:
    # in a `rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` run of this command it
    # calibrates the counters (bytes counted per byte copied)
    if args.calib_mb > 0:
        nb = args.calib_mb << 20
        ca = L.DeviceArray(np.zeros(nb // 8))
        cb = L.DeviceArray(np.zeros(nb // 8))
        L.device_copy(cb.ptr, ca.ptr, nb)
        L.device_sync()
        ca.close(); cb.close()
        log(f"[bench] calibration copy {nb} B")

    # ---- reference fp64 solution (CPU baseline leg doubles as the checker) ----
    cpu = cpu_all = cpu_mixed = None
    xref = None
    if not args.no_cpu:
        kind = "reference" if O.ref_available() else "port"
        which = "ref64" if kind == "reference" else "oracle"
        times = []
        for r in range(max(1, args.cpu_sample)):
            t1 = time.perf_counter()
            rc, xr, ln, un = O.cpu_lu_ref(c.n, c.row, c.col, c.val, q, c.rhs, which=which)
            times.append(time.perf_counter() - t1)
            log(f"[bench] cpu {kind} factor+solve run {r}: {times[-1]:.2f}s rc={rc}")
            xref = xr
        cpu = {"value": round(1.0 / float(np.median(times)), 4), "unit": "refactor+solve/s",
               "cores": 1, "kind": kind,
               "sample": f"{len(times)} full fp64 factor+solve runs of the n=99906 matrix "
                         f"(splu lu_factor + lu_solve_perm, same q), median, 1 thread, "
                         f"{cpu_model()}"}
        P = args.cpu_workers or host_cpus()
        rate, wall = cpu_all_cores(O, c, q, which, P)
        log(f"[bench] cpu {kind} on {P} threads: {rate:.2f} solves/s ({wall:.1f}s)")
        cpu_all = {"value": round(rate, 3), "unit": "refactor+solve/s", "cores": P, "kind": kind,
                   "sample": f"{P} threads x 2 independent fp64 factor+solve runs of the n=99906 "
                             f"matrix at once, solves over wall time; {cpu_model()}, "
                             f"{os.cpu_count()} CPUs online, {P} in this process's share"}
        cpu_mixed = cpu_mixed_c0(L, O)
        log(f"[bench] cpu mixed path c0: {cpu_mixed['value']} /s err={cpu_mixed['rel_err']:.2e}")

    def run_steps(prec, k, BB=None):
        BB = BB or B
        lat = []
        reps = []
        for _ in range(k):
            t1 = time.perf_counter()
            rr = BB.refactor(prec)
            sr = BB.solve_device()
            lat.append((time.perf_counter() - t1) * 1e3)
            reps.append((rr, sr))
        return lat, reps

    def measure(prec, BB=None):
        BB = BB or B
        run_steps(prec, args.warmup, BB)
        L.device_sync()
        BB.kernel_stats(reset=True)
        t1 = time.perf_counter()
        lat, reps = run_steps(prec, args.steps, BB)
        L.device_sync()
        wall = time.perf_counter() - t1
        st = BB.kernel_stats()
        x = BB.get_x()
        err = float(np.linalg.norm(x - xref) / np.linalg.norm(xref)) if xref is not None else None
        return dict(wall=wall, lat=lat, reps=reps, stats=st, err=err)

    log("[bench] fp64 path ...")
    m64 = measure("fp64")
    log(f"[bench] fp64 path: {args.steps / m64['wall']:.2f} steps/s err={m64['err']}")
    log("[bench] mixed path ...")
    m32 = measure("fp32")
    log(f"[bench] mixed path: {args.steps / m32['wall']:.2f} steps/s err={m32['err']}")

    # the same mixed step at the IR tolerance of the reference's selector
    # (src/selector.c:124-126: 1e-9) instead of its test driver's 1e-10; the
    # headline value above keeps 1e-10
    log("[bench] mixed path at the selector's IR tolerance 1e-9 ...")
    B9 = L.Batch(circ, ir_inner_tol=1e-9, ir_outer_tol=1e-9)
    B9.set_rhs(c.rhs)
    m9 = measure("fp32", B9)
    B9.close()
    sel_tol = {"ir_tol": 1e-9, "source": "selector.c:124-126",
               "value": round(args.steps / m9["wall"], 3),
               "ms_per_step": round(1e3 * m9["wall"] / args.steps, 3), "rel_err": m9["err"],
               "ir_inner_iters": int(np.median([r[1]["inner_total"] for r in m9["reps"]])),
               "pass_1e-6": bool(all(r[1]["rc"] == 0 for r in m9["reps"]) and
                                 (m9["err"] is None or m9["err"] <= 1e-6))}
    log(f"[bench] selector tolerance: {sel_tol['value']:.2f} steps/s err={m9['err']}")

    # the mixed step with the refactor on fp64 columns and the fp32 factors the
    # solves read rounded from them (lsp_options.refactor_fp64); fp32 storage
    # for the solves, fp64 for the factorization's own columns
    log("[bench] mixed path, refactor on fp64 columns ...")
    Bc = L.Batch(circ, refactor_fp64=1)
    Bc.set_rhs(c.rhs)
    mc = measure("fp32", Bc)
    Bc.close()
    fp64_cols = {"value": round(args.steps / mc["wall"], 3),
                 "ms_per_step": round(1e3 * mc["wall"] / args.steps, 3), "rel_err": mc["err"],
                 "p50_ms": round(float(np.percentile(mc["lat"], 50)), 3),
                 "p99_ms": round(float(np.percentile(mc["lat"], 99)), 3),
                 "ir_inner_iters": int(np.median([r[1]["inner_total"] for r in mc["reps"]])),
                 "pass_1e-6": bool(all(r[1]["rc"] == 0 for r in mc["reps"]) and
                                   (mc["err"] is None or mc["err"] <= 1e-6))}
    log(f"[bench] refactor on fp64 columns: {fp64_cols['value']:.2f} steps/s err={mc['err']}")

    # the north-star comparison: the batched resident mixed path against the
    # plain one-matrix-per-launch fp64 path on the Newton mix (512 circuits,
    # n in [100, 5000], seed 3, 20 steps)
    newton = None
    if args.newton_circuits > 0:
        from luspice_amd import bench_newton
        B.close()
        newton = bench_newton.compare(ncirc=args.newton_circuits, steps=args.newton_steps)

    st = m32["stats"]
    n, nnz = c.n, c.nnz
    fac_b, tri_b, spmv_b = algorithmic_bytes(inf, nnz, n, 4)
    rf_ms = st["refactor_ms"] / max(1, st["refactor_count"])
    tri_ms = (st["fwd_ms"] + st["bwd_ms"]) / max(1, st["tri_count"])
    # dominant kernel: the single launch with the most GPU time per step.  The
    # refactor kernel runs once per step; each solve direction is one sparse
    # launch plus the dense-block launches, timed together (fwd_ms, bwd_ms),
    # so the larger direction bounds any single solve launch from above.
    per_step_fac = st["refactor_ms"] / args.steps
    per_step_dir = max(st["fwd_ms"], st["bwd_ms"]) / args.steps
    if per_step_fac >= per_step_dir:
        dom, ach_bytes, dur = "lsp::refactor_kernel<float>", fac_b, rf_ms
    else:
        dom, ach_bytes, dur = "trisolve (fwd+bwd launches)", tri_b, tri_ms
    achieved = ach_bytes / (dur * 1e-3) / 1e9 if dur > 0 else 0.0
    lat = np.array(m32["lat"])
    value = args.steps / m32["wall"]
    inner = [r[1]["inner_total"] for r in m32["reps"]]
    ok = all(r[1]["rc"] == 0 for r in m32["reps"]) and (m32["err"] is None or m32["err"] <= 1e-6)
    line = {
        "metric": "refactor+solve/sec at rel err <=1e-6, mixed vs fp64 GPU path; p50/p99 latency",
        "value": round(value, 3),
        "unit": "refactor+solve/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * m32["wall"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (fp32 factor storage)",
        "data": "synthetic",
        "config": {"workload": "c1 power-grid MNA 316x316 + 50 pads, n=99906, seed 1, "
                               "one refactor + fp64 GMRES-IR solve per step",
                   "n": n, "nnz": nnz, "lnnz": inf["lnnz"], "unnz": inf["unnz"],
                   "levels": inf["nlevels"], "parallelism": "single GPU"},
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "rel_err_vs_fp64_cpu": m32["err"],
        "pass_1e-6": bool(ok),
        "ir_inner_iters": int(np.median(inner)) if inner else None,
        "fp64_path": {"value": round(args.steps / m64["wall"], 3),
                      "ms_per_step": round(1e3 * m64["wall"] / args.steps, 3),
                      "rel_err": m64["err"]},
        "mixed_over_fp64": round(value / (args.steps / m64["wall"]), 3),
        "mixed_at_selector_tol": sel_tol,
        "mixed_refactor_fp64_columns": fp64_cols,
        "kernel_ms": {"refactor_kernel": round(rf_ms, 4),
                      "dense_block": round(st["dense_ms"] / max(1, st["refactor_count"]), 4),
                      "trisolve_pair": round(tri_ms, 4),
                      "trisolves_per_step": round(st["tri_count"] / args.steps, 2)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(dom),
                     "algorithmic_bytes": int(ach_bytes), "duration_ms": round(dur, 4)},
        "roofline_other": [tri_roofline(tri_b, tri_ms)],
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "cpu_mixed_c0": cpu_mixed,
        "batched_newton": newton,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
