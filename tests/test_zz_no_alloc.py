"""trisolve_f32_launch_noalloc (lu_gpu.h): the fp32-factor solve of the
reference's mixed path with its workspace in the library's static per-GPU
array instead of caller buffers.  Called through the C ABI, its x must equal
trisolve_f32_launch's bit for bit (same kernels), on two calls in a row --
the config-0 factors (n = 1020, the reference's fp32 splu factors from the
golden fixture), then a smaller circuit (n = 300) -- for the <<<1,1>>> and
<<<1,256>>> forms; a size beyond LU_NOALLOC_MAX_N is rejected before any
launch."""
import ctypes as C
import os

import numpy as np
import pytest

import luspice_amd as L
from luspice_amd import _native as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
vp = C.c_void_p
MAX_N = 262144          # LU_NOALLOC_MAX_N


def _sig(name, *args):
    f = getattr(N.lib, name)
    f.restype = C.c_int
    f.argtypes = list(args)
    return f


_TS = [C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
ts_f32 = _sig("trisolve_f32_launch", *_TS, vp, vp, C.c_int, vp)
ts_noalloc = _sig("trisolve_f32_launch_noalloc", *_TS, C.c_int, vp)


def _dev(a):
    return L.DeviceArray(np.ascontiguousarray(a))


def _inputs_c0():
    g = np.load(os.path.join(GOLD, "c0.npz"))
    F = {k: g["f32_" + k] for k in ("pinv", "Lp", "Li", "Lx", "Up", "Ui", "Ux")}
    return int(g["n"]), F, g["q"], g["rhs"]


def _inputs_small():
    c = L.mna("sized", n=300, seed=77)
    q = L.colamd(c)
    F = O.lu_factor(c.n, c.row, c.col, c.val, q)
    F = {k: (F[k].astype(np.float32) if k in ("Lx", "Ux") else F[k])
         for k in ("pinv", "Lp", "Li", "Lx", "Up", "Ui", "Ux")}
    return c.n, F, q, c.rhs


def _solve(n, F, q, rhs, parallel, noalloc):
    d = {k: _dev(v) for k, v in F.items()}
    dq, db, dx = _dev(np.asarray(q, np.int32)), _dev(rhs), _dev(np.full(n, np.nan))
    args = [n, d["Lp"].ptr, d["Li"].ptr, d["Lx"].ptr, d["Up"].ptr, d["Ui"].ptr, d["Ux"].ptr,
            d["pinv"].ptr, dq.ptr, db.ptr, dx.ptr]
    if noalloc:
        rc = ts_noalloc(*args, parallel, None)
    else:
        wi, wd = _dev(np.zeros(n, np.int32)), _dev(np.zeros(2 * n))
        rc = ts_f32(*args, wi.ptr, wd.ptr, parallel, None)
    assert rc == 0
    L.device_sync()
    return dx.get()


def test_noalloc_trisolve_equals_allocating_entry_point():
    for name, make in (("c0 n=1020", _inputs_c0), ("sized n=300", _inputs_small)):
        n, F, q, rhs = make()
        for parallel in (0, 1):
            print(f"case: noalloc trisolve {name} parallel={parallel}", flush=True)
            want = _solve(n, F, q, rhs, parallel, noalloc=False)
            got = _solve(n, F, q, rhs, parallel, noalloc=True)
            assert np.all(np.isfinite(want))
            assert np.array_equal(got, want), name
    # and both equal the oracle's solve with the same fp32 factors (fp64
    # accumulation, the same column order)
    n, F, q, rhs = _inputs_c0()
    x = _solve(n, F, q, rhs, 1, noalloc=True)
    assert O.rel_l2(x, O.trisolve_f32(n, F, q, rhs)) <= 1e-12


def test_noalloc_rejects_sizes_beyond_its_array():
    print("case: noalloc size check", flush=True)
    args = [MAX_N + 1] + [None] * 10
    assert ts_noalloc(*args, 0, None) == 2
    args[0] = -1
    assert ts_noalloc(*args, 0, None) == 2
