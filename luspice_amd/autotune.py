"""Autotuner for the batch options that have no closed-form best value.

Knobs (all fields of lsp_options):
  sell_sigma        SELL-C-64 sorting window (0 = whole matrix); rows are
                    sorted by length inside each window, so a small window
                    keeps x locality, a large one removes padding
  refactor_blocks   persistent refactor grid (0 = twice the widest level)
  trisolve_blocks   persistent trisolve grid (0 = twice the widest level)
  lds_cap           fp64 LDS doubles per refactor wave (columns above spill
                    to a global scratch slice)
  tail_density      minimum pattern density of the dense trailing block
                    (lower = larger dense block)

A candidate is timed on the real step: set values, refactor (fp32), GMRES-IR
solve, for ``steps`` steps after one warm-up; each configuration must also
keep the solution within ``tol`` of the default configuration's result.  The
search is coordinate-wise (one knob at a time, starting from the defaults),
which stays within a few dozen candidates.  Results are cached in a JSON file
keyed by a structural signature of the batch (circuit count, total n, nnz,
factor nonzeros, levels), so a workload class is tuned once.

    from luspice_amd import autotune
    best, table = autotune.tune(circuits)          # dict of option overrides
    B = luspice_amd.Batch(circuits, **best)
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import Batch

SPACE = {
    "sell_sigma": [0, 64, 256, 1024, 4096],
    "refactor_blocks": [0, 64, 128, 256, 512, 1024],
    "trisolve_blocks": [0, 64, 128, 256, 512, 1024],
    "lds_cap": [512, 1024, 2048],
    "tail_density": [0.3, 0.4, 0.5, 0.6, 0.7],
}

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "luspice_amd", "autotune.json")


def signature(B: Batch) -> str:
    infos = [B.info(i) for i in range(len(B.circuits))]
    return "c{}_n{}_z{}_f{}_l{}".format(len(infos), sum(i["n"] for i in infos), sum(i["nnz"] for i in infos),
                                      sum(i["lnnz"] + i["unnz"] for i in infos),
                                      max(i["nlevels"] for i in infos))


def _measure(circuits, opts: dict, steps: int, vals, rhs):
    B = Batch(circuits, **opts)
    try:
        B.set_values(vals)
        B.refactor("fp32")
        x, _ = B.solve(rhs)                        # warm-up
        t = []
        for _ in range(steps):
            t0 = time.perf_counter()
            B.set_values(vals)
            rr = B.refactor("fp32")
            x, sr = B.solve(rhs)
            t.append(time.perf_counter() - t0)
            if rr["rc"] != 0 or sr["rc"] != 0:
                return None, x
        return float(np.median(t)), x
    finally:
        B.close()


def tune(circuits, steps: int = 3, tol: float = 1e-8, space: dict | None = None,
         cache: str | None = DEFAULT_CACHE, log=None) -> tuple[dict, list]:
    """Returns (best option overrides, [(opts, seconds)] of every candidate)."""
    space = space or SPACE
    vals = np.concatenate([c.val for c in circuits])
    rhs = np.concatenate([c.rhs for c in circuits])
    probe = Batch(circuits, upload=False)
    key = signature(probe)
    probe.close()
    if cache and os.path.exists(cache):
        with open(cache) as f:
            stored = json.load(f)
        if key in stored:
            return stored[key], []
    best = {}
    t_best, x_ref = _measure(circuits, best, steps, vals, rhs)
    table = [(dict(best), t_best)]
    for knob, values in space.items():
        for v in values:
            cand = dict(best)
            cand[knob] = v
            if cand == best:
                continue
            t, x = _measure(circuits, cand, steps, vals, rhs)
            table.append((cand, t))
            if log:
                log(f"[autotune] {cand}: {t}")
            if t is None or np.linalg.norm(x - x_ref) > tol * np.linalg.norm(x_ref):
                continue
            if t < t_best:
                best, t_best = cand, t
    if cache:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        stored = {}
        if os.path.exists(cache):
            with open(cache) as f:
                stored = json.load(f)
        stored[key] = best
        with open(cache, "w") as f:
            json.dump(stored, f, indent=1)
    return best, table
