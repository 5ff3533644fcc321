"""Batched Newton driver and the serving loop.

GPU tests: every result is compared with the fp64 oracle (rel err <= 1e-6),
a request's per-step solutions are the same alone and in a concurrent mix,
and admission never holds more estimated device memory than the budget.
CPU test: the request generator is seeded and Poisson-like.
"""
import numpy as np
import pytest

import luspice_amd as L
from luspice_amd import serve as S
from luspice_amd import workloads as W
from oracle import oracle as O


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_request_generator_seeded():
    a = S.make_requests(12, rate=500.0, nmin=200, nmax=2000, seed=4)
    b = S.make_requests(12, rate=500.0, nmin=200, nmax=2000, seed=4)
    assert [r.circuit.n for r in a] == [r.circuit.n for r in b]
    assert [r.arrival for r in a] == [r.arrival for r in b]
    assert all(200 <= r.circuit.n <= 2000 for r in a)
    assert all(5 <= r.steps <= 50 for r in a)
    assert np.all(np.diff([r.arrival for r in a]) >= 0) and a[0].arrival == 0.0
    va, ba = a[3].step_inputs(2)
    vb, bb = b[3].step_inputs(2)
    assert np.array_equal(va, vb) and np.array_equal(ba, bb)


@pytest.mark.gpu
def test_newton_batched_matches_oracle():
    print("case: 24-circuit Newton batch, 3 steps", flush=True)
    circ = W.newton_mix(24, 100, 3000, seed=3)
    NB = W.NewtonBatched(circ, 3, seed=3)
    aoff = np.cumsum([0] + [c.nnz for c in circ])
    voff = np.cumsum([0] + [c.n for c in circ])
    for s in range(3):
        rr, sr = NB.run_step(s)
        assert rr["rc"] == 0 and sr["rc"] == 0, (rr, sr)
        x = NB.solution()
        vals, rhs = NB.host[s]
        for i, c in enumerate(circ):
            rc, xr, _, _ = O.cpu_lu_ref(c.n, c.row, c.col, vals[aoff[i]:aoff[i + 1]], c.q,
                                         rhs[voff[i]:voff[i + 1]])
            assert _rel(x[voff[i]:voff[i + 1]], xr) <= 1e-6, (s, i, c.n)


@pytest.mark.gpu
@pytest.mark.parametrize("opts", [{}, {"refactor_fp64": 1}])
def test_graph_step_matches_eager(opts):
    print("case: 12-circuit Newton batch, graph replay vs eager, 4 steps", opts, flush=True)
    circ = W.newton_mix(12, 100, 2000, seed=5)
    NB = W.NewtonBatched(circ, 4, seed=5, **opts)
    for s in range(4):
        rr, sr = NB.run_step(s)                      # eager, with its host synchronisation
        assert rr["rc"] == 0 and sr["rc"] == 0, (rr, sr)
        xe = NB.solution().copy()
        rg, sg = NB.run_step(s, graph=True)          # graph replay of the same step
        assert rg["rc"] == 0 and sg["rc"] == 0, (rg, sg)
        xg = NB.solution()
        assert sg["converged"] == len(circ)
        assert _rel(xg, xe) <= 1e-13, s
    assert NB.B.kernel_stats()["graph_steps"] == 4


@pytest.mark.gpu
def test_serve_same_alone_and_mixed():
    print("case: 6 concurrent requests vs alone", flush=True)
    reqs = S.make_requests(6, rate=1e6, nmin=200, nmax=3000, smin=3, smax=4, seed=4)
    srv = S.Server(8 << 30, slots=6, keep_results=True)
    st = srv.run(reqs)
    srv.close()
    assert st["errors"] == 0 and st["solves"] == sum(r.steps for r in reqs), st
    for r in reqs:
        alone = S.Request(r.rid, 0.0, r.circuit, r.steps, r.seed)
        s1 = S.Server(8 << 30, slots=1, keep_results=True)
        s1.run([alone])
        s1.close()
        assert len(alone.results) == r.steps
        for a, b in zip(r.results, alone.results):
            assert _rel(a, b) <= 1e-12
        vals, rhs = r.step_inputs(r.steps - 1)
        c = r.circuit
        rc, xr, _, _ = O.cpu_lu_ref(c.n, c.row, c.col, vals, c.q, rhs)
        assert _rel(r.results[-1], xr) <= 1e-6


@pytest.mark.gpu
def test_serve_cohorts_merge_and_overlap():
    """Continuous batching: three requests prepared together run as one cohort
    (one launch per step for all three); a fourth that arrives while that
    cohort is on the GPU is admitted as a cohort of its own and its steps
    overlap the first cohort's in time.  Each result equals the request run
    alone to 1e-12 and the fp64 CPU LU to 1e-6."""
    print("case: cohorts merge and overlap", flush=True)
    reqs = S.make_requests(4, rate=1e6, nmin=500, nmax=1900, smin=40, smax=40, seed=6)
    for r in reqs[:3]:
        r.arrival = 0.0
    reqs[3].arrival = 0.02
    srv = S.Server(8 << 30, slots=4, keep_results=True, gather_ms=200.0)
    srv.log_steps = True
    st = srv.run(reqs)
    srv.close()
    print(st, flush=True)
    assert st["errors"] == 0 and st["solves"] == 160, st
    by_cohort = {}
    for cid, ids, t0, t1 in srv.step_log:
        by_cohort.setdefault(cid, []).append((ids, t0, t1))
    first = [c for c, steps in by_cohort.items() if any(len(ids) == 3 for ids, _, _ in steps)]
    late = [c for c, steps in by_cohort.items() if all(ids == (reqs[3].rid,) for ids, _, _ in steps)]
    assert first and late, by_cohort.keys()
    a, b = by_cohort[first[0]], by_cohort[late[0]]
    assert b[0][1] < a[-1][2]                 # the late cohort started before the first one ended
    for r in reqs:
        alone = S.Request(r.rid, 0.0, r.circuit, r.steps, r.seed)
        s1 = S.Server(8 << 30, slots=1, keep_results=True)
        s1.run([alone])
        s1.close()
        for x, y in zip(r.results, alone.results):
            assert _rel(x, y) <= 1e-12
        vals, rhs = r.step_inputs(r.steps - 1)
        c = r.circuit
        rc, xr, _, _ = O.cpu_lu_ref(c.n, c.row, c.col, vals, c.q, rhs)
        assert _rel(r.results[-1], xr) <= 1e-6


@pytest.mark.gpu
def test_serve_admission_respects_budget():
    print("case: admission under a tight HBM budget", flush=True)
    reqs = S.make_requests(5, rate=1e6, nmin=1000, nmax=2000, smin=2, smax=3, seed=5)
    est = []
    for r in reqs:
        c = r.circuit
        B = L.Batch(L.Circuit(c.n, c.row, c.col, c.val, c.rhs, q=c.q), upload=False)
        est.append(B.estimate_bytes())
        B.close()
    budget = int(max(est) * 1.5)              # never room for two of the larger ones
    srv = S.Server(budget, slots=5)
    st = srv.run(reqs)
    srv.close()
    assert st["errors"] == 0 and st["solves"] == sum(r.steps for r in reqs)
    assert st["peak_admitted_bytes"] <= budget


@pytest.mark.gpu
def test_autotune_small_space(c0, c0_q):
    print("case: autotune on c0, reduced space", flush=True)
    from luspice_amd import autotune
    circ = L.Circuit(c0.n, c0.row, c0.col, c0.val, c0.rhs, q=c0_q)
    best, table = autotune.tune([circ], steps=1, cache=None,
                                space={"sell_sigma": [0, 256], "tail_density": [0.3, 0.6]})
    # the defaults, then each listed value of each knob (starting from the best so far)
    assert len(table) == 5 and all(t is not None for _, t in table)
    assert set(best) <= {"sell_sigma", "tail_density"}
    B = L.Batch(circ, **best)
    B.refactor("fp32")
    x, rep = B.solve(c0.rhs)
    rc, xr, _, _ = O.cpu_lu_ref(c0.n, c0.row, c0.col, c0.val, c0_q, c0.rhs)
    assert _rel(x, xr) <= 1e-6
