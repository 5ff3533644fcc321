"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where built, to
the reference's own C (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; never by luspice_amd/.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

ip = C.POINTER(C.c_int)
dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)
lp = C.POINTER(C.c_long)


def _p(a, t=None):
    if a is None:
        return None
    if t is None:
        t = {np.dtype(np.int32): C.c_int, np.dtype(np.float64): C.c_double,
             np.dtype(np.float32): C.c_float}[a.dtype]
    return a.ctypes.data_as(C.POINTER(t))


def build_oracle():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        L = C.CDLL(ORACLE_SO)
        L.orc_lu_factor_export.restype = C.c_int
        L.orc_lu_factor_export.argtypes = [C.c_int, ip, ip, dp, ip, C.c_double, ip, ip, ip, dp,
                                           C.c_long, ip, ip, dp, C.c_long, lp, lp]
        L.orc_cpu_lu_ref.restype = C.c_int
        L.orc_cpu_lu_ref.argtypes = [C.c_int, ip, ip, dp, ip, dp, C.c_double, dp, lp, lp]
        L.orc_level_compute.argtypes = [C.c_int, ip, ip, ip, ip]
        L.orc_level_hist.argtypes = [C.c_int, ip, C.c_int, ip]
        L.orc_residual.restype = C.c_double
        L.orc_residual.argtypes = [C.c_int, C.c_long, ip, ip, dp, dp, dp]
        L.orc_rel_l2.restype = C.c_double
        L.orc_rel_l2.argtypes = [C.c_int, dp, dp]
        L.orc_csr_spmv.argtypes = [C.c_int, ip, ip, dp, dp, dp]
        L.orc_trisolve_f32.argtypes = [C.c_int, ip, ip, fp, ip, ip, fp, ip, ip, dp, dp]
        L.orc_gmres_ir.restype = C.c_int
        L.orc_gmres_ir.argtypes = [C.c_int, C.c_long, ip, ip, dp, dp, ip, ip, fp, ip, ip, fp,
                                   ip, ip, C.c_int, C.c_int, C.c_double, C.c_double, dp, ip, ip,
                                   dp]
        L.orc_ruiz.argtypes = [C.c_int, C.c_long, ip, ip, dp, C.c_int, dp, dp]
        L.orc_diag_shift.argtypes = [C.c_int, C.c_long, ip, ip, dp, C.c_double, dp]
        L.orc_sell_order.restype = C.c_long
        L.orc_sell_order.argtypes = [C.c_int, ip, C.c_int, ip, ip, ip]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    """The reference's own C compiled from its sources (oracle/Makefile ref)."""
    global _ref
    if _ref is None:
        if not ref_available():
            return None
        L = C.CDLL(REF_SO)
        for nm, vt in (("ref64", dp), ("ref32", fp)):
            f = getattr(L, nm + "_lu_factor_export")
            f.restype = C.c_int
            f.argtypes = [C.c_int, ip, ip, vt, ip, C.c_double, ip, ip, ip, vt, C.c_long, ip, ip,
                          vt, C.c_long, lp, lp]
            g = getattr(L, nm + "_factor_solve")
            g.restype = C.c_int
            g.argtypes = [C.c_int, ip, ip, vt, ip, C.c_double, vt, vt, lp, lp]
        L.cpu_lu_ref.restype = C.c_int
        L.cpu_lu_ref.argtypes = [C.c_int, ip, ip, dp, ip, dp, C.c_double, dp, lp, lp]
        L.dsh_level_compute.argtypes = [C.c_int, ip, ip, ip, ip]
        L.dsh_level_hist.argtypes = [C.c_int, ip, C.c_int, ip]
        _ref = L
    return _ref


def _csc_q(n, row, col, val, q):
    """q-order CSC, rows in COO order (src/mp/mp_lu_bench.cu:150-186)."""
    qq = np.arange(n) if q is None else np.asarray(q)
    invq = np.empty(n, np.int64)
    invq[qq] = np.arange(n)
    cols = invq[col]
    order = np.argsort(cols, kind="stable")
    Ap = np.zeros(n + 1, np.int64)
    np.add.at(Ap, cols + 1, 1)
    Ap = np.cumsum(Ap).astype(np.int32)
    return Ap, np.ascontiguousarray(row[order], np.int32), np.ascontiguousarray(val[order])


def natural_csc(n, row, col, val):
    return _csc_q(n, np.asarray(row, np.int32), np.asarray(col, np.int32),
                  np.asarray(val, np.float64), None)


def lu_factor(n, row, col, val, q, tol=1.0, which="oracle"):
    """Factor P*A(:,q) = L*U.  which: 'oracle' (this repository's C
    restatement), 'ref64' / 'ref32' (the reference's splu.h, fp64 / fp32).
    Returns dict with pinv, Lp, Li, Lx, Up, Ui, Ux (splu storage order)."""
    Ap, Ai, Ax = natural_csc(n, row, col, val)
    qa = None if q is None else np.ascontiguousarray(q, np.int32)
    if which == "oracle":
        f, vdt = lib().orc_lu_factor_export, np.float64
    else:
        f, vdt = getattr(ref(), which + "_lu_factor_export"), (np.float32 if which == "ref32" else np.float64)
    Axv = np.ascontiguousarray(Ax, vdt)
    ln, un = C.c_long(), C.c_long()
    cap = 1
    dummy_i = np.zeros(1, np.int32)
    dummy_v = np.zeros(1, vdt)
    pinv = np.empty(n, np.int32)
    Lp = np.empty(n + 1, np.int32)
    Up = np.empty(n + 1, np.int32)
    rc = f(n, _p(Ap), _p(Ai), _p(Axv), _p(qa), tol, _p(pinv), _p(Lp), _p(dummy_i), _p(dummy_v),
           0, _p(Up), _p(dummy_i), _p(dummy_v), 0, C.byref(ln), C.byref(un))
    if rc >= 0:
        return {"singular": rc}
    Li = np.empty(ln.value, np.int32)
    Lx = np.empty(ln.value, vdt)
    Ui = np.empty(un.value, np.int32)
    Ux = np.empty(un.value, vdt)
    rc = f(n, _p(Ap), _p(Ai), _p(Axv), _p(qa), tol, _p(pinv), _p(Lp), _p(Li), _p(Lx),
           ln.value, _p(Up), _p(Ui), _p(Ux), un.value, C.byref(ln), C.byref(un))
    assert rc == -1, rc
    del cap
    return dict(pinv=pinv, Lp=Lp, Li=Li, Lx=Lx, Up=Up, Ui=Ui, Ux=Ux, singular=-1)


def cpu_lu_ref(n, row, col, val, q, b, tol=1.0, which="oracle"):
    """x of the cpu_lu_ref chain (factor + lu_solve_perm)."""
    Ap, Ai, Ax = natural_csc(n, row, col, val)
    qa = None if q is None else np.ascontiguousarray(q, np.int32)
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty(n)
    ln, un = C.c_long(), C.c_long()
    if which == "oracle":
        rc = lib().orc_cpu_lu_ref(n, _p(Ap), _p(Ai), _p(Ax), _p(qa), _p(b), tol, _p(x),
                                  C.byref(ln), C.byref(un))
    elif which == "ref_cpu_lu_ref":
        rc = ref().cpu_lu_ref(n, _p(Ap), _p(Ai), _p(Ax), _p(qa), _p(b), tol, _p(x),
                              C.byref(ln), C.byref(un))
    else:
        rc = ref().ref64_factor_solve(n, _p(Ap), _p(Ai), _p(Ax), _p(qa), tol, _p(b), _p(x),
                                      C.byref(ln), C.byref(un))
    return rc, x, ln.value, un.value


def levels(n, Up, Ui, which="oracle"):
    lv = np.empty(n, np.int32)
    nl = C.c_int()
    L = lib() if which == "oracle" else ref()
    fn = L.orc_level_compute if which == "oracle" else L.dsh_level_compute
    fn(n, _p(np.ascontiguousarray(Up, np.int32)), _p(np.ascontiguousarray(Ui, np.int32)),
       _p(lv), C.byref(nl))
    return lv, nl.value


def level_hist(level, nlevels, which="oracle"):
    """hist[l] = number of columns at level l (src/level.c:43-48)."""
    lv = np.ascontiguousarray(level, np.int32)
    hist = np.empty(max(int(nlevels), 1), np.int32)
    L = lib() if which == "oracle" else ref()
    fn = L.orc_level_hist if which == "oracle" else L.dsh_level_hist
    fn(len(lv), _p(lv), int(nlevels), _p(hist))
    return hist[:int(nlevels)]


def residual(n, row, col, val, b, x):
    r1 = np.ascontiguousarray(np.asarray(row) + 1, np.int32)
    c1 = np.ascontiguousarray(np.asarray(col) + 1, np.int32)
    return lib().orc_residual(n, len(val), _p(r1), _p(c1), _p(np.ascontiguousarray(val, np.float64)),
                              _p(np.ascontiguousarray(b, np.float64)),
                              _p(np.ascontiguousarray(x, np.float64)))


def rel_l2(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return lib().orc_rel_l2(len(a), _p(a), _p(b))


def csr_spmv(n, rowptr, colidx, val, x):
    y = np.empty(n)
    lib().orc_csr_spmv(n, _p(np.ascontiguousarray(rowptr, np.int32)),
                       _p(np.ascontiguousarray(colidx, np.int32)),
                       _p(np.ascontiguousarray(val, np.float64)),
                       _p(np.ascontiguousarray(x, np.float64)), _p(y))
    return y


def trisolve_f32(n, F, q, b):
    x = np.empty(n)
    qa = None if q is None else np.ascontiguousarray(q, np.int32)
    lib().orc_trisolve_f32(n, _p(F["Lp"]), _p(F["Li"]), _p(np.ascontiguousarray(F["Lx"], np.float32)),
                           _p(F["Up"]), _p(F["Ui"]), _p(np.ascontiguousarray(F["Ux"], np.float32)),
                           _p(F["pinv"]), _p(qa), _p(np.ascontiguousarray(b, np.float64)), _p(x))
    return x


def gmres_ir(n, row, col, val, b, F32, q, max_outer=7, m=50, inner_tol=1e-10, outer_tol=1e-10):
    """The reference's mixed solve (fp32 factors, fp64 GMRES-IR), restated."""
    r1 = np.ascontiguousarray(np.asarray(row) + 1, np.int32)
    c1 = np.ascontiguousarray(np.asarray(col) + 1, np.int32)
    x = np.empty(n)
    ti, od = C.c_int(), C.c_int()
    fr = C.c_double()
    qa = None if q is None else np.ascontiguousarray(q, np.int32)
    rc = lib().orc_gmres_ir(n, len(val), _p(r1), _p(c1), _p(np.ascontiguousarray(val, np.float64)),
                            _p(np.ascontiguousarray(b, np.float64)), _p(F32["Lp"]), _p(F32["Li"]),
                            _p(np.ascontiguousarray(F32["Lx"], np.float32)), _p(F32["Up"]),
                            _p(F32["Ui"]), _p(np.ascontiguousarray(F32["Ux"], np.float32)),
                            _p(F32["pinv"]), _p(qa), max_outer, m, inner_tol, outer_tol, _p(x),
                            C.byref(ti), C.byref(od), C.byref(fr))
    return x, dict(rc=rc, inner=ti.value, outer=od.value, res=fr.value)


def diag_shift(n, row, col, val, shift):
    """Factor-only diagonal shift (orc_diag_shift): the first (i, i) entry of
    each row, in entry order, becomes a + shift*|a|."""
    r1 = np.ascontiguousarray(np.asarray(row) + 1, np.int32)
    c1 = np.ascontiguousarray(np.asarray(col) + 1, np.int32)
    v = np.ascontiguousarray(val, np.float64)
    out = np.empty_like(v)
    lib().orc_diag_shift(n, len(v), _p(r1), _p(c1), _p(v), shift, _p(out))
    return out


def ruiz(n, row, col, val, sweeps=30):
    r1 = np.ascontiguousarray(np.asarray(row) + 1, np.int32)
    c1 = np.ascontiguousarray(np.asarray(col) + 1, np.int32)
    v = np.array(val, np.float64)
    R = np.empty(n)
    Cc = np.empty(n)
    lib().orc_ruiz(n, len(v), _p(r1), _p(c1), _p(v), sweeps, _p(R), _p(Cc))
    return v, R, Cc


def sell_order(n, rowptr, C_=32):
    order = np.empty(n, np.int32)
    ns = (n + C_ - 1) // C_
    lens = np.empty(ns, np.int32)
    offs = np.empty(ns + 1, np.int32)
    tot = lib().orc_sell_order(n, _p(np.ascontiguousarray(rowptr, np.int32)), C_, _p(order),
                               _p(lens), _p(offs))
    return order, lens, offs, tot
