"""Synthetic MNA workloads and Newton-step drivers.

Circuits come from the host generator (csrc/host/mna.cpp):
  * ``newton_mix`` builds the batched Newton re-factorisation workload:
    circuits with n log-uniform in [nmin, nmax] and a fixed pattern each;
  * every Newton step multiplies each value by lognormal(0, sigma) and draws a
    fresh right-hand side uniform in [-1, 1] (``perturb``).

Two drivers run the same steps:
  * ``NewtonBatched`` -- all circuits in one device-resident batch.  Per step it
    does one value upload from HBM, one refactor launch and one batched
    GMRES-IR solve.  The step inputs are staged in device memory before the
    timed loop.
  * ``NewtonPlain``   -- the plain fp64 path: one batch per circuit, and
    every step uploads values and right-hand side from the host, refactors
    in fp64, solves and copies x back.
"""
from __future__ import annotations

import math

import numpy as np

from . import Batch, Circuit, DeviceArray, colamd, mna, perturb


def newton_mix(ncirc: int = 512, nmin: int = 100, nmax: int = 5000, seed: int = 3,
               lr: float = 0.02) -> list[Circuit]:
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(math.log(nmin), math.log(nmax), ncirc))).astype(int)
    out = []
    for i, n in enumerate(ns):
        c = mna("sized", n=int(n), lr=lr, seed=seed * 1_000_003 + i)
        c.q = colamd(c)
        out.append(c)
    return out


def step_inputs(circuits, step: int, seed: int = 3, sigma: float = 0.1):
    """Concatenated values (batch value order) and right-hand sides for one step."""
    vals, rhs = [], []
    for i, c in enumerate(circuits):
        v, b = perturb(seed * 1_000_003 + i, step, c.val, c.n, sigma)
        vals.append(v)
        rhs.append(b)
    return np.concatenate(vals), np.concatenate(rhs)


class NewtonBatched:
    """All circuits in one batch; step inputs resident in HBM."""

    def __init__(self, circuits, steps: int, seed: int = 3, sigma: float = 0.1, host_inputs=None, **opts):
        self.circuits = circuits
        self.B = Batch(circuits, **opts)
        self.steps = steps
        self.host = host_inputs if host_inputs is not None else [
            step_inputs(circuits, s, seed, sigma) for s in range(steps)]
        self.dev_vals = [DeviceArray(v) for v, _ in self.host]
        self.dev_rhs = [DeviceArray(b) for _, b in self.host]

    def run_step(self, s: int, prec: str = "fp32", graph: bool = False):
        """Values and b from HBM, refactor + GMRES-IR; graph=True replays the
        step as one HIP graph (recorded on first use)."""
        B = self.B
        B.set_values_device(self.dev_vals[s].ptr)
        B.set_rhs(self.dev_rhs[s].ptr)
        if graph:
            return B.graph_step(prec)
        rr = B.refactor(prec)
        sr = B.solve_device()
        return rr, sr

    def solution(self) -> np.ndarray:
        return self.B.get_x()

    def close(self):
        for a in self.dev_vals + self.dev_rhs:
            a.close()
        self.dev_vals, self.dev_rhs = [], []
        self.B.close()


class NewtonPlain:
    """One batch per circuit, fp64 factors, host inputs every step."""

    def __init__(self, circuits, steps: int, seed: int = 3, sigma: float = 0.1, host_inputs=None, **opts):
        self.circuits = circuits
        self.Bs = [Batch(c, **opts) for c in circuits]
        self.steps = steps
        self.host = host_inputs if host_inputs is not None else [
            step_inputs(circuits, s, seed, sigma) for s in range(steps)]
        self.aoff = np.cumsum([0] + [c.nnz for c in circuits])
        self.voff = np.cumsum([0] + [c.n for c in circuits])

    def run_step(self, s: int, prec: str = "fp64"):
        vals, rhs = self.host[s]
        xs, reps = [], []
        for i, B in enumerate(self.Bs):
            B.set_values(vals[self.aoff[i]:self.aoff[i + 1]])
            rr = B.refactor(prec)
            x, sr = B.solve(rhs[self.voff[i]:self.voff[i + 1]])
            xs.append(x)
            reps.append((rr, sr))
        return np.concatenate(xs), reps

    def close(self):
        for B in self.Bs:
            B.close()
        self.Bs = []
