"""The drop-in boundary through the reference's own entry points, on config 0
(n = 1020) and on Newton steps of small circuits.

* spmv.h (src/spmv.h:34-57): read_s0_to_csr / read_rhs / csr_free on an s0
  file, spmv_host and csr_alloc_device + spmv_launch, each equal to the CSR
  product of the oracle bit for bit (same per-row order, contraction off).
* lu_gpu.h / lu_gpu_par.h launchers: the <<<1,1>>> and <<<1,256>>> fp64
  factors equal the compiled reference's splu factors (tests/golden/c0.npz)
  bit for bit, and their solves give its x; the fp32 factors (fp32
  arithmetic, as the reference kernel) agree with splu built with LU_FLOAT
  to fp32 rounding.
* lu_static.h: the reference's caller flow (bench/test_lu.c:136-209):
  symbolic once, numeric twice with the symbolic count staying 1,
  factor_cache read as splu's lu_factors, levels of its U, trisolve with a
  residual at fp64 level, then dsh_gmres_ir; then Newton steps whose P must
  equal splu's P for each step's values.
* Batches with repivot = 3 reproduce splu's P at every Newton step; with the
  default repivot = 1 a step whose fixed pivot is zero gets new pivots and
  solves in the same call.
"""
import ctypes as C
import os

import numpy as np
import pytest

import luspice_amd as L
from luspice_amd import _native as N
from luspice_amd import s0tools as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ip, dp, fp, vp = C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_float), C.c_void_p


class CSRMatrix(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_longlong), ("rowptr", ip), ("colidx", ip),
                ("values", dp)]


class LuCsc(C.Structure):        # splu.h:29-34
    _fields_ = [("n", C.c_int), ("Lp", ip), ("Li", ip), ("Lx", dp), ("Up", ip), ("Ui", ip),
                ("Ux", dp), ("Lcap", C.c_long), ("Ucap", C.c_long), ("lnnz", C.c_long),
                ("unnz", C.c_long)]


class LuFactors(C.Structure):    # splu.h:36-41
    _fields_ = [("n", C.c_int), ("pinv", ip), ("L", LuCsc), ("U", LuCsc), ("singular_col", C.c_int)]


class DshLuStatic(C.Structure):  # lu_static.h
    _fields_ = [("n", C.c_int), ("cnnz", C.c_long), ("lnnz", C.c_long), ("unnz", C.c_long),
                ("cap", C.c_long), ("Ap", ip), ("Ai", ip), ("Ax", fp), ("P", ip), ("Q", ip),
                ("min_diag_U", C.c_double), ("pivot_growth", C.c_double),
                ("dev_allocated", C.c_int), ("factor_cache", vp), ("pin", C.c_double)]


def _sig(name, res, *args):
    f = getattr(N.lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("read_s0_to_csr", C.c_int, C.c_char_p, C.POINTER(CSRMatrix))
_sig("csr_free", None, C.POINTER(CSRMatrix))
_sig("read_rhs", C.c_int, C.c_char_p, C.c_int, dp)
_sig("spmv_host", C.c_int, C.POINTER(CSRMatrix), dp, dp, dp)
_sig("csr_alloc_device", C.c_int, C.POINTER(CSRMatrix), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp))
_sig("spmv_launch", None, C.c_int, vp, vp, vp, vp, vp, vp)
_LU = [C.c_int, vp, vp, vp, vp, vp, vp, vp, C.c_long, vp, vp, vp, C.c_long, vp, vp, vp, vp, vp, vp, vp]
_sig("lu_factor_launch", C.c_int, *_LU)
_sig("lu_factor_f32_launch", C.c_int, *_LU)
_sig("lu_factor_par_launch", C.c_int, *(_LU[:15] + _LU[16:]))
_TS = [C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
_sig("trisolve_launch", C.c_int, *_TS, vp)
_sig("trisolve_par_launch", C.c_int, *_TS, vp)
_sig("trisolve_f32_launch", C.c_int, *_TS, C.c_int, vp)
_sig("dsh_lu_static_init", C.c_int, C.POINTER(DshLuStatic))
_sig("dsh_lu_static_free", None, C.POINTER(DshLuStatic))
_sig("dsh_symbolic", C.c_int, C.c_int, C.c_long, ip, ip, ip, C.POINTER(DshLuStatic))
_sig("dsh_numeric", C.c_int, C.POINTER(DshLuStatic), ip, ip, dp, C.c_int)
_sig("dsh_trisolve", C.c_int, C.POINTER(DshLuStatic), dp, dp)
_sig("dsh_gmres_ir", C.c_int, C.POINTER(DshLuStatic), dp, dp, C.c_int, C.c_int, C.c_double, dp, ip)
_sig("dsh_symbolic_count", C.c_long)
_sig("dsh_lu_status_str", C.c_char_p, C.c_int)


@pytest.fixture(scope="module")
def g0():
    return dict(np.load(os.path.join(GOLD, "c0.npz")))


def _dev(a):
    return L.DeviceArray(np.ascontiguousarray(a))


def _csr_of(n, row, col, val):
    order = np.lexsort((np.arange(len(row)), row))
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, row + 1, 1)
    return np.cumsum(rp).astype(np.int32), col[order].astype(np.int32), val[order]


# --------------------------------------------------------------------------
def test_spmv_h_entry_points(tmp_path, g0):
    print("case: spmv.h entry points on c0", flush=True)
    n = int(g0["n"])
    row, col, val, rhs = g0["row"], g0["col"], g0["val"], g0["rhs"]
    mtx, rhsf = S.write_s0(str(tmp_path / "c0"), n, row, col, val, rhs)
    A = CSRMatrix()
    assert N.lib.read_s0_to_csr(mtx.encode(), C.byref(A)) == 0
    assert A.n == n and A.nnz == len(val)
    rp = np.ctypeslib.as_array(A.rowptr, (n + 1,)).copy()
    ci = np.ctypeslib.as_array(A.colidx, (len(val),)).copy()
    vv = np.ctypeslib.as_array(A.values, (len(val),)).copy()
    rp0, ci0, vv0 = _csr_of(n, row, col, val)
    assert np.array_equal(rp, rp0) and np.array_equal(ci, ci0) and np.array_equal(vv, vv0)
    b = np.empty(n)
    assert N.lib.read_rhs(rhsf.encode(), n, N.ptr(b)) == 0
    assert np.array_equal(b, rhs)
    x = np.random.default_rng(8).uniform(-1, 1, n)
    want = O.csr_spmv(n, rp, ci, vv, x)
    y = np.empty(n)
    t = C.c_double(-1.0)
    assert N.lib.spmv_host(C.byref(A), N.ptr(x), N.ptr(y), C.byref(t)) == 0
    assert np.array_equal(y, want) and t.value > 0
    d_rp, d_ci, d_v = vp(), vp(), vp()
    assert N.lib.csr_alloc_device(C.byref(A), C.byref(d_rp), C.byref(d_ci), C.byref(d_v)) == 0
    dx, dy = _dev(x), _dev(np.zeros(n))
    N.lib.spmv_launch(n, d_rp, d_ci, d_v, dx.ptr, dy.ptr, None)
    L.device_sync()
    assert np.array_equal(dy.get(), want)
    for p in (d_rp, d_ci, d_v):
        N.lib.lsp_dev_free(p)
    N.lib.csr_free(C.byref(A))
    assert not A.rowptr and not A.colidx and not A.values


# --------------------------------------------------------------------------
def _launch_factor(fn, n, Ap, Ai, Ax, q, lcap, ucap, vdt, par=False):
    bufs = dict(Ap=_dev(Ap), Ai=_dev(Ai), Ax=_dev(Ax.astype(vdt)), diag=_dev(q),
                Lp=_dev(np.zeros(n + 1, np.int32)), Li=_dev(np.zeros(lcap, np.int32)),
                Lx=_dev(np.zeros(lcap, vdt)), Up=_dev(np.zeros(n + 1, np.int32)),
                Ui=_dev(np.zeros(ucap, np.int32)), Ux=_dev(np.zeros(ucap, vdt)),
                pinv=_dev(np.zeros(n, np.int32)), xi=_dev(np.zeros(n + 4 * 1024 + 2 * n + 1, np.int32)),
                x=_dev(np.zeros(n, vdt)), ln=_dev(np.zeros(1, np.int32)), un=_dev(np.zeros(1, np.int32)),
                sing=_dev(np.full(1, -7, np.int32)))

    b = bufs
    args = [n, b["Ap"].ptr, b["Ai"].ptr, b["Ax"].ptr, b["diag"].ptr, b["Lp"].ptr, b["Li"].ptr,
            b["Lx"].ptr, lcap, b["Up"].ptr, b["Ui"].ptr, b["Ux"].ptr, ucap, b["pinv"].ptr, b["xi"].ptr]
    if not par:
        args.append(b["x"].ptr)
    args += [b["ln"].ptr, b["un"].ptr, b["sing"].ptr, None]
    assert fn(*args) == 0
    L.device_sync()
    ln, un = int(b["ln"].get()[0]), int(b["un"].get()[0])
    out = dict(singular=int(b["sing"].get()[0]), pinv=b["pinv"].get(), Lp=b["Lp"].get(),
               Li=b["Li"].get()[:ln], Lx=b["Lx"].get()[:ln], Up=b["Up"].get(), Ui=b["Ui"].get()[:un],
               Ux=b["Ux"].get()[:un])
    return out, bufs


def _launch_solve(fn, n, bufs, q, rhs, *extra):
    bx, dq, dxo = _dev(rhs), _dev(q), _dev(np.zeros(n))
    wi, wd = _dev(np.zeros(n, np.int32)), _dev(np.zeros(2 * n))
    assert fn(n, bufs["Lp"].ptr, bufs["Li"].ptr, bufs["Lx"].ptr, bufs["Up"].ptr, bufs["Ui"].ptr,
              bufs["Ux"].ptr, bufs["pinv"].ptr, dq.ptr, bx.ptr, dxo.ptr, wi.ptr, wd.ptr, *extra, None) == 0
    L.device_sync()
    return dxo.get()


def test_lu_gpu_launchers_match_compiled_reference(g0):
    n = int(g0["n"])
    c = L.Circuit(n, g0["row"], g0["col"], g0["val"], g0["rhs"])
    q = g0["q"]
    Ap, Ai, Ax = c.csc(q)                       # q-order CSC with original row ids
    ln, un = int(g0["lnnz"]), int(g0["unnz"])
    cap_l, cap_u = ln + 64, un + 64
    keys = ("pinv", "Lp", "Li", "Lx", "Up", "Ui", "Ux")
    for name, fn, par in (("lu_factor_launch <<<1,1>>>", N.lib.lu_factor_launch, False),
                          ("lu_factor_par_launch <<<1,256>>>", N.lib.lu_factor_par_launch, True)):
        print("case:", name, flush=True)
        F, bufs = _launch_factor(fn, n, Ap, Ai, Ax, q, cap_l, cap_u, np.float64, par)
        assert F["singular"] == -1
        for k in keys:
            assert np.array_equal(F[k], g0[f"f64_{k}"]), (name, k)
        for sname, sfn in (("trisolve_launch", N.lib.trisolve_launch),
                           ("trisolve_par_launch", N.lib.trisolve_par_launch)):
            x = _launch_solve(sfn, n, bufs, q, g0["rhs"])
            print("  ", sname, "rel err", O.rel_l2(x, g0["x"]), flush=True)
            assert O.rel_l2(x, g0["x"]) <= 1e-14, sname
    print("case: lu_factor_f32_launch <<<1,1>>>", flush=True)
    F, bufs = _launch_factor(N.lib.lu_factor_f32_launch, n, Ap, Ai, Ax, q, cap_l, cap_u, np.float32)
    assert F["singular"] == -1
    for k in ("pinv", "Lp", "Li", "Up", "Ui"):
        assert np.array_equal(F[k], g0[f"f32_{k}"]), k
    for k in ("Lx", "Ux"):
        ref = g0[f"f32_{k}"].astype(np.float64)
        d = np.abs(F[k].astype(np.float64) - ref)
        assert np.all(d <= 1e-4 * np.abs(ref) + 1e-6 * np.abs(ref).max()), k
    for par in (0, 1):
        x = _launch_solve(N.lib.trisolve_f32_launch, n, bufs, q, g0["rhs"], par)
        assert O.rel_l2(x, g0["x"]) <= 1e-3


# --------------------------------------------------------------------------
def _coo1(c):
    return (np.ascontiguousarray(c.row + 1, np.int32), np.ascontiguousarray(c.col + 1, np.int32))


def _residual(c, x, b):
    return O.residual(c.n, c.row, c.col, c.val, b, x)


def test_lu_static_reference_flow(g0):
    """bench/test_lu.c:136-209 restated: symbolic once, numeric twice (count
    stays), factor_cache as splu's lu_factors, levels of its U, trisolve with
    ||b - A x|| / ||b|| <= 1e-6 (PASS line), then GMRES-IR."""
    print("case: lu_static flow on c0", flush=True)
    n = int(g0["n"])
    c = L.Circuit(n, g0["row"], g0["col"], g0["val"], g0["rhs"])
    r1, c1 = _coo1(c)
    F = DshLuStatic()
    assert N.lib.dsh_lu_static_init(C.byref(F)) == 0
    F.pin = 0.0
    q = np.ascontiguousarray(g0["q"], np.int32)
    cnt0 = N.lib.dsh_symbolic_count()
    assert N.lib.dsh_symbolic(n, c.nnz, N.ptr(r1), N.ptr(c1), N.ptr(q), C.byref(F)) == 0
    assert N.lib.dsh_symbolic_count() == cnt0 + 1
    assert N.lib.dsh_numeric(C.byref(F), N.ptr(r1), N.ptr(c1), N.ptr(c.val), c.nnz) == 0
    assert F.lnnz == g0["lnnz"] and F.unnz == g0["unnz"] and F.cap == F.lnnz + F.unnz
    Ud = g0["f64_Ux"][g0["f64_Up"][1:] - 1]
    assert abs(F.min_diag_U - np.abs(Ud[Ud != 0]).min()) <= 1e-12 * np.abs(Ud).max()
    assert N.lib.dsh_numeric(C.byref(F), N.ptr(r1), N.ptr(c1), N.ptr(c.val), c.nnz) == 0
    assert N.lib.dsh_symbolic_count() == cnt0 + 1            # the symbolic phase was reused
    assert np.array_equal(np.ctypeslib.as_array(F.P, (n,)), g0["f64_pinv"])
    # the caller reads the factors through factor_cache as splu's lu_factors
    fac = C.cast(F.factor_cache, C.POINTER(LuFactors)).contents
    assert fac.n == n and fac.L.lnnz == g0["lnnz"] and fac.U.unnz == g0["unnz"]
    assert np.array_equal(np.ctypeslib.as_array(fac.pinv, (n,)), g0["f64_pinv"])
    Up = np.ctypeslib.as_array(fac.U.Up, (n + 1,)).copy()
    Lp = np.ctypeslib.as_array(fac.L.Lp, (n + 1,)).copy()
    assert np.array_equal(Up, g0["f64_Up"]) and np.array_equal(Lp, g0["f64_Lp"])
    Ui = np.ctypeslib.as_array(fac.U.Ui, (int(Up[-1]),)).copy()
    Ux = np.ctypeslib.as_array(fac.U.Ux, (int(Up[-1]),)).copy()
    Li = np.ctypeslib.as_array(fac.L.Li, (int(Lp[-1]),)).copy()
    for k in range(n):
        a, b = slice(Up[k], Up[k + 1]), slice(Lp[k], Lp[k + 1])
        assert Ui[Up[k + 1] - 1] == k and Li[Lp[k]] == k       # U diagonal last, L unit first
        assert np.array_equal(np.sort(Ui[a]), np.sort(g0["f64_Ui"][a]))
        assert np.array_equal(np.sort(Li[b]), np.sort(g0["f64_Li"][b]))
    np.testing.assert_allclose(Ux[Up[1:] - 1], Ud, rtol=1e-12)
    lv, nl = O.levels(n, Up, Ui, which="oracle")
    assert nl == g0["nlevels"] and np.array_equal(lv, g0["level"])
    x = np.empty(n)
    assert N.lib.dsh_trisolve(C.byref(F), N.ptr(c.rhs), N.ptr(x)) == 0
    res = _residual(c, x, c.rhs)
    print("  trisolve residual", res, flush=True)
    assert res <= 1e-6 and res <= 1e-12                      # PASS, at fp64 level
    assert O.rel_l2(x, g0["x"]) <= 1e-12
    seq = np.zeros(8)
    it = C.c_int()
    xg = np.empty(n)
    assert N.lib.dsh_gmres_ir(C.byref(F), N.ptr(c.rhs), N.ptr(xg), 7, 50, 1e-10, N.ptr(seq), C.byref(it)) == 0
    print("  gmres-ir", it.value, seq[:3], flush=True)
    assert O.rel_l2(xg, g0["x"]) <= 1e-6 and it.value >= 1 and seq[0] > seq[1]
    # a trisolve after GMRES-IR is back on the fp64 factors
    assert N.lib.dsh_trisolve(C.byref(F), N.ptr(c.rhs), N.ptr(x)) == 0
    assert O.rel_l2(x, g0["x"]) <= 1e-12
    assert N.lib.dsh_lu_status_str(1) == b"SINGULAR"
    N.lib.dsh_lu_static_free(C.byref(F))
    assert not F.factor_cache


def test_lu_static_newton_steps_follow_splu_pivots():
    """Each dsh_numeric of a Newton step gives splu's P for that step's values
    (about a quarter of the pivot rows change per step on this circuit), and
    dsh_trisolve the oracle's solution."""
    c = L.mna("sized", n=600, seed=41)
    q = L.colamd(c)
    r1, c1 = _coo1(c)
    F = DshLuStatic()
    N.lib.dsh_lu_static_init(C.byref(F))
    assert N.lib.dsh_symbolic(c.n, c.nnz, N.ptr(r1), N.ptr(c1), N.ptr(q), C.byref(F)) == 0
    P0 = None
    changed = 0
    for step in range(5):
        print("case: lu_static Newton step", step, flush=True)
        v, b = (c.val, c.rhs) if step == 0 else L.perturb(7, step, c.val, c.n)
        assert N.lib.dsh_numeric(C.byref(F), N.ptr(r1), N.ptr(c1), N.ptr(v), c.nnz) == 0
        R = O.lu_factor(c.n, c.row, c.col, v, q)
        P = np.ctypeslib.as_array(F.P, (c.n,)).copy()
        assert np.array_equal(P, R["pinv"]), step
        assert F.lnnz == len(R["Li"]) and F.unnz == len(R["Ui"])
        if P0 is not None:
            changed += int((P != P0).sum())
        P0 = P
        x = np.empty(c.n)
        assert N.lib.dsh_trisolve(C.byref(F), N.ptr(b), N.ptr(x)) == 0
        rc, xo, _, _ = O.cpu_lu_ref(c.n, c.row, c.col, v, q, b)
        assert O.rel_l2(x, xo) <= 1e-10
    assert changed > 0                                       # re-pivoting was exercised
    N.lib.dsh_lu_static_free(C.byref(F))


def test_batch_repivot_exact_follows_splu_each_step():
    """A 3-circuit batch with repivot = 3 and fp64 factors: after every Newton
    step each circuit's P equals the oracle's splu P for that step's values,
    and the solutions match the oracle."""
    cs = [L.mna("sized", n=n, seed=60 + i) for i, n in enumerate((300, 900, 150))]
    qs = [L.colamd(c) for c in cs]
    B = L.Batch([L.Circuit(c.n, c.row, c.col, c.val, c.rhs, q=q) for c, q in zip(cs, qs)], repivot=3)
    total = 0
    for step in range(1, 4):
        print("case: batch repivot=3 step", step, flush=True)
        vb = [L.perturb(11, step, c.val, c.n) for c in cs]
        B.set_values(np.concatenate([v for v, _ in vb]))
        rep = B.refactor("fp64")
        assert rep["rc"] == 0, rep
        total += rep["repivoted"]
        x, sr = B.solve(np.concatenate([b for _, b in vb]))
        assert sr["rc"] == 0, sr
        for i, (c, q) in enumerate(zip(cs, qs)):
            v, b = vb[i]
            R = O.lu_factor(c.n, c.row, c.col, v, q)
            assert np.array_equal(B.export_factors(i)["pinv"], R["pinv"]), (step, i)
            rc, xo, _, _ = O.cpu_lu_ref(c.n, c.row, c.col, v, q, b)
            assert O.rel_l2(x[B.voff[i]:B.voff[i + 1]], xo) <= 1e-10
    assert total > 0


def test_zero_pivot_gets_new_pivots_in_the_same_call():
    """The default repivot = 1: the step whose fixed pivot is exactly zero is
    refactored with the pivots splu picks for it, inside the same call, and
    solves to the dense solution."""
    print("case: zero pivot, repivot=1", flush=True)
    row = np.array([0, 0, 1, 1, 1, 2, 2], np.int32)
    col = np.array([0, 1, 0, 1, 2, 1, 2], np.int32)
    v0 = np.array([4.0, 1.0, 1.0, 4.0, 1.0, 1.0, 4.0])
    b = np.array([1.0, 2.0, 3.0])
    q = np.arange(3, dtype=np.int32)
    B = L.Batch(L.Circuit(3, row, col, v0, b, q=q), tail_max=-1)
    v1 = np.array([2.0, 1.0, 1.0, 0.5, 1.0, 1.0, 4.0])
    B.set_values(v1)
    rep = B.refactor("fp64")
    print(rep, flush=True)
    assert rep["rc"] == 0 and rep["repivoted"] == 1 and rep["singular_circuits"] == 0
    R = O.lu_factor(3, row, col, v1, q)
    assert np.array_equal(B.export_factors(0)["pinv"], R["pinv"])
    x, sr = B.solve(b)
    A = np.zeros((3, 3))
    A[row, col] = v1
    assert sr["rc"] == 0 and O.rel_l2(x, np.linalg.solve(A, b)) <= 1e-12


def test_small_pivot_repivots_only_from_mode_2():
    """A step whose fixed pivot is tiny but nonzero (|U(1,1)| = 1e-5 against
    |L(2,1)| = 1, failing refactor_ptol = 1e-3): the default repivot = 1 keeps
    the pivots and reports the bad pivot (GMRES-IR still solves it); repivot =
    2 takes the pivots splu picks for the step within the same call."""
    print("case: small pivot, repivot=1 and 2", flush=True)
    row = np.array([0, 0, 1, 1, 1, 2, 2], np.int32)
    col = np.array([0, 1, 0, 1, 2, 1, 2], np.int32)
    v0 = np.array([4.0, 1.0, 1.0, 4.0, 1.0, 1.0, 4.0])
    b = np.array([1.0, 2.0, 3.0])
    q = np.arange(3, dtype=np.int32)
    v1 = np.array([2.0, 1.0, 1.0, 0.5 + 1e-5, 1.0, 1.0, 4.0])
    A = np.zeros((3, 3))
    A[row, col] = v1
    for mode, repiv in ((1, 0), (2, 1)):
        B = L.Batch(L.Circuit(3, row, col, v0, b, q=q), tail_max=-1, repivot=mode)
        B.set_values(v1)
        rep = B.refactor("fp64")
        print(mode, rep, flush=True)
        assert rep["rc"] == 0 and rep["repivoted"] == repiv, rep
        assert rep["bad_pivots"] == (1 if mode == 1 else 0), rep
        if mode == 2:
            assert np.array_equal(B.export_factors(0)["pinv"], O.lu_factor(3, row, col, v1, q)["pinv"])
        x, sr = B.solve(b)
        assert sr["rc"] == 0 and O.rel_l2(x, np.linalg.solve(A, b)) <= 1e-10, sr
        B.close()
