"""Fused iterative-refinement step (one launch per Richardson step for a whole
batch) against the direct solve and the CPU oracle.

Tolerances: one step from x = 0 equals the direct solve Q U^-1 L^-1 P b up to
fp64 summation order (1e-12 relative for fp64 factors, 1e-9 for fp32 factors,
whose products are accumulated in fp64 on both paths); a step from any x equals
x + trisolve(b - A x) within the same bounds; the reported ||b - A x|| / ||b||
within 1e-12 of numpy's; solutions through ir_method=1 within 1e-6 relative L2
of the fp64 oracle (1e-9 with fp64 factors).
"""
import numpy as np
import pytest

import luspice_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _ax(c, x):
    y = np.zeros(c.n)
    np.add.at(y, c.row, c.val * x[c.col])
    return y


def _oracle_x(c, q):
    rc, x, ln, un = O.cpu_lu_ref(c.n, c.row, c.col, c.val, q, c.rhs)
    assert rc == -1
    return x


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-9)])
def test_ir_step_from_zero_is_the_direct_solve(c0, c0_q, prec, tol):
    print("case: c0 ir_step from 0,", prec, flush=True)
    B = L.Batch(L.Circuit(c0.n, c0.row, c0.col, c0.val, c0.rhs, q=c0_q))
    assert B.refactor(prec)["rc"] == 0
    x1, rel = B.ir_step(np.zeros(c0.n), c0.rhs)
    xd = B.trisolve(c0.rhs)
    assert O.rel_l2(x1, xd) <= tol
    assert abs(rel[0] - 1.0) <= 1e-15


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-9)])
def test_ir_step_is_one_richardson_step(c0, c0_q, prec, tol):
    """x1 = x0 + M^-1 (b - A x0): the residual (fp64 values, original row
    order), the permuted solve and the update of the original-order x."""
    print("case: c0 ir_step from a random x,", prec, flush=True)
    B = L.Batch(L.Circuit(c0.n, c0.row, c0.col, c0.val, c0.rhs, q=c0_q))
    assert B.refactor(prec)["rc"] == 0
    x0 = np.random.default_rng(3).uniform(-1, 1, c0.n)
    x1, rel = B.ir_step(x0, c0.rhs)
    r0 = c0.rhs - _ax(c0, x0)
    want = x0 + B.trisolve(r0)
    assert O.rel_l2(x1, want) <= tol
    assert abs(rel[0] - np.linalg.norm(r0) / np.linalg.norm(c0.rhs)) <= 1e-12 * rel[0]


def test_ir_method_fused_fp32_c0_reaches_oracle(c0, c0_q):
    print("case: c0 fp32 ir_method=1", flush=True)
    B = L.Batch(L.Circuit(c0.n, c0.row, c0.col, c0.val, c0.rhs, q=c0_q), ir_method=1)
    assert B.refactor("fp32")["rc"] == 0
    x, sr = B.solve(c0.rhs)
    err = O.rel_l2(x, _oracle_x(c0, c0_q))
    print("fused fp32 rel err", err, sr, flush=True)
    assert sr["rc"] == 0 and err <= 1e-6
    assert sr["outer"] == 1 and 1 <= sr["inner_total"] <= 12     # the fused steps alone converged


def test_ir_method_fused_batch_with_dense_tails_and_global_work_vector():
    """Uneven batch: circuits with dense trailing blocks, one above the 8192
    rows whose work vector fits in LDS; fp64 factors converge in two steps,
    fp32 factors within the fused step limit."""
    cs = [L.mna("sized", n=n, seed=40 + i) for i, n in enumerate([12000, 300, 3000, 77, 1500])]
    cs.append(L.mna("c2", n=3000, seed=5))
    for c in cs:
        c.q = L.colamd(c)
    print("case: fused batch n =", [c.n for c in cs], flush=True)
    rhs = np.concatenate([c.rhs for c in cs])
    xo = [_oracle_x(c, c.q) for c in cs]
    for prec, tol in (("fp64", 1e-9), ("fp32", 1e-6)):
        B = L.Batch(cs, ir_method=1)
        assert any(B.info(i)["tail_T"] > 0 for i in range(len(cs)))
        assert B.refactor(prec)["rc"] == 0
        x, sr = B.solve(rhs)
        print(prec, sr, flush=True)
        assert sr["rc"] == 0 and sr["converged"] == len(cs)
        if prec == "fp64":
            assert sr["inner_total"] <= 3
        for i, c in enumerate(cs):
            assert O.rel_l2(x[B.voff[i]:B.voff[i + 1]], xo[i]) <= tol, (prec, i)
        B.close()


def test_ir_method_fused_hands_unconverged_circuits_to_gmres(c0, c0_q):
    """One fused step cannot reach 1e-10 with fp32 factors: GMRES-IR continues
    from the fused iterate and the answer still meets the pass line."""
    print("case: c0 fp32 one fused step then GMRES-IR", flush=True)
    B = L.Batch(L.Circuit(c0.n, c0.row, c0.col, c0.val, c0.rhs, q=c0_q), ir_method=1,
                ir_fused_steps=1)
    assert B.refactor("fp32")["rc"] == 0
    x, sr = B.solve(c0.rhs)
    err = O.rel_l2(x, _oracle_x(c0, c0_q))
    print("fused 1 + gmres rel err", err, sr, flush=True)
    assert sr["rc"] == 0 and err <= 1e-6 and sr["outer"] >= 2


def test_ir_method_fused_respects_the_active_mask():
    cs = [L.mna("sized", n=n, seed=70 + i) for i, n in enumerate([400, 900, 250])]
    print("case: fused with an active mask", flush=True)
    B = L.Batch(cs, ir_method=1)
    assert B.refactor("fp32")["rc"] == 0
    rhs = np.concatenate([c.rhs for c in cs])
    x_all, sr = B.solve(rhs)
    assert sr["rc"] == 0
    B.set_active([True, False, True])
    rhs2 = rhs * 2.0
    x2, sr2 = B.solve(rhs2)
    assert sr2["rc"] == 0 and sr2["converged"] == 2
    for i in (0, 2):
        a, b = B.voff[i], B.voff[i + 1]
        assert O.rel_l2(x2[a:b], 2.0 * x_all[a:b]) <= 1e-8


def test_fused_graph_step_matches_eager():
    """The Newton step graph with ir_method=1 records refactor + the fused
    steps; every replay equals the eager step (same launches per circuit; a
    converged circuit's later launches return at once) and the fp64 CPU LU."""
    from luspice_amd import workloads as W
    print("case: 12-circuit Newton batch, fused graph replay vs eager, 4 steps", flush=True)
    circ = W.newton_mix(12, 100, 2000, seed=5)
    NB = W.NewtonBatched(circ, 4, seed=5, ir_method=1)
    aoff = np.cumsum([0] + [c.nnz for c in circ])
    voff = np.cumsum([0] + [c.n for c in circ])
    for s in range(4):
        rr, sr = NB.run_step(s)
        assert rr["rc"] == 0 and sr["rc"] == 0, (rr, sr)
        xe = NB.solution().copy()
        rg, sg = NB.run_step(s, graph=True)
        assert rg["rc"] == 0 and sg["rc"] == 0, (rg, sg)
        assert sg["converged"] == len(circ), sg
        xg = NB.solution()
        assert O.rel_l2(xg, xe) <= 1e-13, s
        vals, rhs = NB.host[s]
        for i, c in enumerate(circ):
            rc, xr, _, _ = O.cpu_lu_ref(c.n, c.row, c.col, vals[aoff[i]:aoff[i + 1]], c.q,
                                         rhs[voff[i]:voff[i + 1]])
            assert O.rel_l2(xg[voff[i]:voff[i + 1]], xr) <= 1e-6, (s, i)
    assert NB.B.kernel_stats()["graph_steps"] == 4
    NB.close()
