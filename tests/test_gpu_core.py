"""Parity of the HIP path against the CPU oracle (config 0 and small batches).

Tolerances: integer structure (P, pattern) bit-exact; fp64 factors within
1e-12 relative (summation order differs from the DFS order); fp32 factor
storage within fp32 rounding of the fp64 factors; SpMV with fp64 values
bitwise (same per-row order, contraction off); GMRES-IR solutions within
1e-6 relative L2 of the fp64 oracle x (the north-star pass line).
"""
import numpy as np
import pytest

import luspice_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _oracle_x(c, q):
    rc, x, ln, un = O.cpu_lu_ref(c.n, c.row, c.col, c.val, q, c.rhs)
    assert rc == -1
    return x


def test_refactor_fp64_matches_oracle_factors(c0, c0_q):
    print("case: c0 refactor fp64", flush=True)
    B = L.Batch(L.Circuit(c0.n, c0.row, c0.col, c0.val, c0.rhs, q=c0_q))
    rep = B.refactor("fp64")
    assert rep["rc"] == 0, rep
    F = B.export_factors(0)
    R = O.lu_factor(c0.n, c0.row, c0.col, c0.val, c0_q)
    assert np.array_equal(F["pinv"], R["pinv"])
    assert np.array_equal(F["Lp"], R["Lp"]) and np.array_equal(F["Up"], R["Up"])
    n = c0.n
    for k in range(n):
        a, b = slice(R["Lp"][k], R["Lp"][k + 1]), slice(R["Up"][k], R["Up"][k + 1])
        oi, oj = np.argsort(R["Li"][a], kind="stable"), np.argsort(R["Ui"][b], kind="stable")
        assert np.array_equal(np.sort(R["Li"][a]), F["Li"][a][np.argsort(F["Li"][a])])
        assert np.array_equal(np.sort(R["Ui"][b]), F["Ui"][b][np.argsort(F["Ui"][b])])
        fl = F["Lx"][a][np.argsort(F["Li"][a], kind="stable")]
        fu = F["Ux"][b][np.argsort(F["Ui"][b], kind="stable")]
        np.testing.assert_allclose(fl, R["Lx"][a][oi], rtol=1e-10, atol=1e-12 * np.abs(R["Lx"][a]).max())
        np.testing.assert_allclose(fu, R["Ux"][b][oj], rtol=1e-10, atol=1e-12 * np.abs(R["Ux"][b]).max())


def test_factor_shift_matches_oracle(c0, c0_q):
    """Factor-only diagonal shift: fp64 factors equal the oracle's factors of
    the shifted copy (a duplicate diagonal entry shows that only the first one
    is shifted), the direct solve inverts the shifted matrix, and GMRES-IR on
    the unshifted values still reaches the fp64 solution."""
    print("case: c0 factor shift", flush=True)
    s = 1e-3
    has = np.zeros(c0.n, bool)
    has[c0.row[c0.row == c0.col]] = True
    i_dup = int(np.flatnonzero(has)[11])
    first = int(np.flatnonzero((c0.row == i_dup) & (c0.col == i_dup))[0])
    row = np.append(c0.row, i_dup).astype(np.int32)
    col = np.append(c0.col, i_dup).astype(np.int32)
    val = np.append(c0.val, 2.0 * c0.val[first])
    B = L.Batch(L.Circuit(c0.n, row, col, val, c0.rhs, q=c0_q), factor_shift=s)
    rep = B.refactor("fp64")
    assert rep["rc"] == 0, rep
    F = B.export_factors(0)
    R = O.lu_factor(c0.n, row, col, O.diag_shift(c0.n, row, col, val, s), c0_q)
    assert np.array_equal(F["pinv"], R["pinv"])
    assert np.array_equal(F["Lp"], R["Lp"]) and np.array_equal(F["Up"], R["Up"])
    for k in range(c0.n):
        a, b = slice(R["Lp"][k], R["Lp"][k + 1]), slice(R["Up"][k], R["Up"][k + 1])
        fl = F["Lx"][a][np.argsort(F["Li"][a], kind="stable")]
        fu = F["Ux"][b][np.argsort(F["Ui"][b], kind="stable")]
        rl = R["Lx"][a][np.argsort(R["Li"][a], kind="stable")]
        ru = R["Ux"][b][np.argsort(R["Ui"][b], kind="stable")]
        np.testing.assert_allclose(fl, rl, rtol=1e-10, atol=1e-12 * np.abs(rl).max())
        np.testing.assert_allclose(fu, ru, rtol=1e-10, atol=1e-12 * np.abs(ru).max())
    # no duplicates from here: the solve of the shifted matrix, then GMRES-IR
    B2 = L.Batch(L.Circuit(c0.n, c0.row, c0.col, c0.val, c0.rhs, q=c0_q), factor_shift=s)
    assert B2.refactor("fp64")["rc"] == 0
    vs = O.diag_shift(c0.n, c0.row, c0.col, c0.val, s)
    v = np.random.default_rng(9).uniform(-1, 1, c0.n)
    As_v = np.zeros(c0.n)
    np.add.at(As_v, c0.row, vs * v[c0.col])
    assert O.rel_l2(B2.trisolve(As_v), v) < 1e-9
    xo = _oracle_x(c0, c0_q)
    assert O.rel_l2(B2.trisolve(c0.rhs), xo) > 1e-6      # the shift took effect
    B3 = L.Batch(L.Circuit(c0.n, c0.row, c0.col, c0.val, c0.rhs, q=c0_q), factor_shift=1e-7)
    assert B3.refactor("fp32")["rc"] == 0
    x, sr = B3.solve(c0.rhs)
    print("shift 1e-7 fp32 GMRES-IR", O.rel_l2(x, xo), sr, flush=True)
    assert sr["rc"] == 0 and O.rel_l2(x, xo) <= 1e-6


def test_trisolve_fp64_direct(c0, c0_q):
    print("case: c0 trisolve fp64", flush=True)
    B = L.Batch(L.Circuit(c0.n, c0.row, c0.col, c0.val, c0.rhs, q=c0_q))
    assert B.refactor("fp64")["rc"] == 0
    x = B.trisolve(c0.rhs)
    xo = _oracle_x(c0, c0_q)
    assert O.rel_l2(x, xo) < 1e-10


def test_spmv_fp64_bitwise(c0):
    print("case: c0 spmv", flush=True)
    B = L.Batch(c0)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, c0.n)
    y = B.spmv(x, "fp64")
    order = np.lexsort((np.arange(c0.nnz), c0.row))
    rp = np.zeros(c0.n + 1, np.int32)
    np.add.at(rp, c0.row + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    yo = O.csr_spmv(c0.n, rp, c0.col[order], c0.val[order], x)
    assert np.array_equal(y, yo)


def test_gmres_ir_fp32_factors_c0(c0, c0_q):
    print("case: c0 gmres-ir fp32", flush=True)
    B = L.Batch(L.Circuit(c0.n, c0.row, c0.col, c0.val, c0.rhs, q=c0_q))
    rep = B.refactor("fp32")
    assert rep["rc"] == 0, rep
    x, sr = B.solve(c0.rhs)
    xo = _oracle_x(c0, c0_q)
    err = O.rel_l2(x, xo)
    print("rel err", err, sr, flush=True)
    assert sr["rc"] == 0 and err <= 1e-6
    # the reference's own mixed path restated on the CPU (fp32 splu factors)
    F32 = O.lu_factor(c0.n, c0.row, c0.col, c0.val, c0_q)
    F32 = dict(F32, Lx=F32["Lx"].astype(np.float32), Ux=F32["Ux"].astype(np.float32))
    xr, info = O.gmres_ir(c0.n, c0.row, c0.col, c0.val, c0.rhs, F32, c0_q)
    assert O.rel_l2(x, xr) <= 1e-6


def test_batch_uneven_same_alone_and_mixed():
    print("case: batch of 6 uneven circuits", flush=True)
    cs = [L.mna("sized", n=n, seed=100 + i) for i, n in enumerate([150, 900, 333, 2100, 77, 1500])]
    B = L.Batch(cs)
    assert B.refactor("fp32")["rc"] == 0
    rhs = np.concatenate([c.rhs for c in cs])
    x, rep = B.solve(rhs)
    assert rep["rc"] == 0, rep
    for i, c in enumerate(cs):
        xi = x[B.voff[i]:B.voff[i + 1]]
        q = L.colamd(c)
        xo = _oracle_x(c, q)
        assert O.rel_l2(xi, xo) <= 1e-6
        solo = L.Batch(c)
        assert solo.refactor("fp32")["rc"] == 0
        xs, r2 = solo.solve(c.rhs)
        assert r2["rc"] == 0
        assert O.rel_l2(xi, xs) <= 1e-6


def test_zero_pivot_reported_then_recovered():
    """With repivot = 0 (pivots kept, failures only reported), a Newton step
    whose values make a fixed pivot exactly zero: the refactor reports the
    circuit as singular (LSP_E_SINGULAR, first bad circuit 0) and
    every kernel returns; the next step's values factor and solve normally.
    The 3 x 3 circuit has no dense trailing block, so the sparse refactor
    kernel takes the zero pivot."""
    print("case: zero pivot then recovery", flush=True)
    tail_max = -1
    row = np.array([0, 0, 1, 1, 1, 2, 2], np.int32)
    col = np.array([0, 1, 0, 1, 2, 1, 2], np.int32)
    v0 = np.array([4.0, 1.0, 1.0, 4.0, 1.0, 1.0, 4.0])
    b = np.array([1.0, 2.0, 3.0])
    B = L.Batch(L.Circuit(3, row, col, v0, b, q=np.arange(3, dtype=np.int32)), tail_max=tail_max,
                repivot=0)
    # U(1,1) = 0.5 - 1 * 1 / 2 = 0 exactly with the fixed diagonal pivots
    B.set_values(np.array([2.0, 1.0, 1.0, 0.5, 1.0, 1.0, 4.0]))
    rep = B.refactor("fp64")
    print(rep, flush=True)
    assert rep["rc"] == -4 and rep["singular_circuits"] == 1 and rep["first_bad_circuit"] == 0
    B.set_values(v0)
    rep = B.refactor("fp64")
    assert rep["rc"] == 0 and rep["singular_circuits"] == 0, rep
    x, sr = B.solve(b)
    A = np.zeros((3, 3))
    A[row, col] = v0
    assert sr["rc"] == 0 and O.rel_l2(x, np.linalg.solve(A, b)) <= 1e-12


def test_ir_stall_falls_back_to_fp64(c0, c0_q):
    """IR with fp32 factors capped at one restart of one inner iteration
    cannot reach the 1e-10 residual on c0: without the fallback the solve
    reports not converged; with it the stalled circuit is refactored in fp64,
    solved again, and lands within 1e-6 of the oracle (one fallback).  A second
    circuit in the same batch with the IR capped the same way shows the
    fallback acts per circuit."""
    print("case: c0 IR stall -> fp64 fallback", flush=True)
    xo = _oracle_x(c0, c0_q)
    circ = L.Circuit(c0.n, c0.row, c0.col, c0.val, c0.rhs, q=c0_q)
    B0 = L.Batch(circ, fallback_fp64=0, ir_outer=1, ir_inner=1)
    assert B0.refactor("fp32")["rc"] == 0
    x0, s0 = B0.solve(c0.rhs)
    print("no fallback", s0, O.rel_l2(x0, xo), flush=True)
    assert s0["rc"] != 0 and s0["converged"] == 0 and s0["fallback"] == 0
    B0.close()
    c1 = L.mna("sized", n=700, seed=41)
    q1 = L.colamd(c1)
    B = L.Batch([circ, L.Circuit(c1.n, c1.row, c1.col, c1.val, c1.rhs, q=q1)], fallback_fp64=1,
                ir_outer=1, ir_inner=1)
    assert B.refactor("fp32")["rc"] == 0
    x, sr = B.solve(np.concatenate([c0.rhs, c1.rhs]))
    print("fallback", sr, flush=True)
    assert sr["rc"] == 0 and sr["fallback"] == 2 and sr["converged"] == 2, sr
    assert O.rel_l2(x[:c0.n], xo) <= 1e-6
    assert O.rel_l2(x[c0.n:], _oracle_x(c1, q1)) <= 1e-6
    B.close()


def test_fp16_refused_without_opt_in(c0, c0_q):
    """FP16 factor storage is off unless lsp_options.allow_fp16 is set: the
    refactor returns LSP_E_ARG before any device work."""
    B = L.Batch(L.Circuit(c0.n, c0.row, c0.col, c0.val, c0.rhs, q=c0_q))
    rep = B.refactor("fp16")
    assert rep["rc"] == -1, rep                 # LSP_E_ARG
    assert B.refactor("fp32")["rc"] == 0
    B.close()


def test_pivot_range_is_the_range_of_u_diagonal(c0, c0_q):
    """lsp_batch_pivot_range (the pivot guard's input) against the diagonal of
    the exported U (its last entry per column), including the dense block's
    columns."""
    print("case: c0 pivot range", flush=True)
    B = L.Batch(L.Circuit(c0.n, c0.row, c0.col, c0.val, c0.rhs, q=c0_q))
    assert B.info(0)["tail_T"] > 0
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-6)):
        assert B.refactor(prec)["rc"] == 0
        F = B.export_factors(0)
        d = np.abs(F["Ux"][F["Up"][1:] - 1])
        lo, hi = B.pivot_range(0)
        np.testing.assert_allclose([lo, hi], [d.min(), d.max()], rtol=tol, err_msg=prec)


def test_refactor_fp64_columns_gives_the_rounded_fp64_factors(c0, c0_q):
    """refactor_fp64: the fp32 factors are fp32 values within one rounding
    (2^-24 relative) of an fp64 refactor's factors (two fp64 refactors may
    differ in the last bits: dependencies are applied in the order their
    flags are seen), and GMRES-IR on them meets the pass line."""
    print("case: c0 fp32 factors rounded from the fp64 refactor", flush=True)
    circ = L.Circuit(c0.n, c0.row, c0.col, c0.val, c0.rhs, q=c0_q)
    B64 = L.Batch(circ)
    assert B64.refactor("fp64")["rc"] == 0
    F64 = B64.export_factors(0)
    B = L.Batch(circ, refactor_fp64=1)
    assert B.refactor("fp32")["rc"] == 0
    F = B.export_factors(0)
    for k in ("pinv", "Lp", "Li", "Up", "Ui"):
        assert np.array_equal(F[k], F64[k]), k
    for k in ("Lx", "Ux"):
        assert np.array_equal(F[k], F[k].astype(np.float32).astype(np.float64)), k
        np.testing.assert_allclose(F[k], F64[k], rtol=2.0 ** -24 * (1 + 1e-6),
                                   atol=1e-13 * np.abs(F64[k]).max(), err_msg=k)
    x, sr = B.solve(c0.rhs)
    err = O.rel_l2(x, _oracle_x(c0, c0_q))
    print("rel err", err, sr, flush=True)
    assert sr["rc"] == 0 and err <= 1e-6
