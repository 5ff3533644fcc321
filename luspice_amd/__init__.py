"""luspice_amd -- MI355X (gfx950) mixed-precision sparse LU for SPICE MNA
matrices: fp32-factor refactorization + fp64 GMRES-IR behind the reference's
C API (lu_gpu.h / lu_gpu_par.h / spmv.h / selector.h / lu_static.h) and a
batched, device-resident C ABI (include/luspice.h).

This package is the Python host mirror of that C ABI; all numerics run in
libluspice_amd.so (HIP kernels + C++ runtime).
"""
from __future__ import annotations

import numpy as np

from . import _native as N
from ._native import check

PREC_FP32, PREC_FP64, PREC_FP16 = 0, 1, 2
_PREC = {"fp32": PREC_FP32, "fp64": PREC_FP64, "fp16": PREC_FP16}


def prec_code(p) -> int:
    return _PREC[p] if isinstance(p, str) else int(p)


class Circuit:
    """An n x n MNA matrix as a 0-based COO pattern plus values and a RHS."""

    def __init__(self, n, row, col, val, rhs=None, q=None):
        self.n = int(n)
        self.row = np.ascontiguousarray(row, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.rhs = None if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64)
        self.q = None if q is None else np.ascontiguousarray(q, dtype=np.int32)

    @property
    def nnz(self):
        return int(self.val.size)

    def dense(self):
        A = np.zeros((self.n, self.n))
        np.add.at(A, (self.row, self.col), self.val)
        return A

    def csc(self, q=None):
        """q-order CSC with original row ids, rows in COO order (build_csc_q)."""
        n = self.n
        qq = np.arange(n, dtype=np.int32) if q is None else np.asarray(q, dtype=np.int32)
        invq = np.empty(n, dtype=np.int64)
        invq[qq] = np.arange(n)
        cols = invq[self.col]
        order = np.argsort(cols, kind="stable")
        Ap = np.zeros(n + 1, dtype=np.int32)
        np.add.at(Ap, cols + 1, 1)
        Ap = np.cumsum(Ap).astype(np.int32)
        return Ap, self.row[order].copy(), self.val[order].copy()


def mna(kind: int | str, n: int = 0, lr: float = 0.02, seed: int = 0) -> Circuit:
    """Synthetic MNA circuit.  kind 0 / 'c0': 25x40 grid + 20 sources (n=1020);
    1 / 'c1': 316x316 power grid + 50 pads (n=99906); 2 / 'sized': n unknowns;
    3 / 'c2': n unknowns, ~6 nonzeros per row, conductances spread over up to
    15 decades (precision sweep)."""
    k = {"c0": 0, "c1": 1, "sized": 2, "c2": 3}.get(kind, kind)
    nn = N.C.c_int()
    nz = N.C.c_long()
    check(N.lib.lsp_mna_generate(k, int(n), float(lr), int(seed), N.C.byref(nn), N.C.byref(nz),
                                 None, None, None, None), "mna size")
    row = np.empty(nz.value, np.int32)
    col = np.empty(nz.value, np.int32)
    val = np.empty(nz.value, np.float64)
    rhs = np.empty(nn.value, np.float64)
    check(N.lib.lsp_mna_generate(k, int(n), float(lr), int(seed), None, None, N.ptr(row),
                                 N.ptr(col), N.ptr(val), N.ptr(rhs)), "mna")
    return Circuit(nn.value, row, col, val, rhs)


def perturb(seed: int, step: int, val0: np.ndarray, n: int, sigma: float = 0.1):
    """Newton-step values val0 * lognormal(0, sigma) and a fresh RHS ~ U[-1,1]."""
    val0 = np.ascontiguousarray(val0, dtype=np.float64)
    val = np.empty_like(val0)
    rhs = np.empty(n, np.float64)
    N.lib.lsp_mna_perturb_values(int(seed), int(step), float(sigma), val0.size, N.ptr(val0),
                                 N.ptr(val), int(n), N.ptr(rhs))
    return val, rhs


class DeviceArray:
    """A numpy array copied into device memory (owned; freed on close)."""

    def __init__(self, host: np.ndarray):
        self.host_dtype = host.dtype
        self.size = host.size
        self.nbytes = host.nbytes
        p = N.C.c_void_p()
        check(N.lib.lsp_dev_alloc(N.C.byref(p), self.nbytes), "dev_alloc")
        self.ptr = p.value
        src = np.ascontiguousarray(host)
        check(N.lib.lsp_dev_copy(N.C.c_void_p(self.ptr), src.ctypes.data_as(N.C.c_void_p), self.nbytes),
              "dev_copy")

    def get(self) -> np.ndarray:
        out = np.empty(self.size, self.host_dtype)
        check(N.lib.lsp_dev_copy(out.ctypes.data_as(N.C.c_void_p), N.C.c_void_p(self.ptr), self.nbytes),
              "dev_copy")
        return out

    def close(self):
        if getattr(self, "ptr", None):
            N.lib.lsp_dev_free(N.C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        self.close()


def device_copy(dst_ptr: int, src_ptr: int, nbytes: int):
    """Device-to-device copy of nbytes (synchronous)."""
    check(N.lib.lsp_dev_copy(N.C.c_void_p(dst_ptr), N.C.c_void_p(src_ptr), int(nbytes)), "dev_copy")


def device_sync():
    check(N.lib.lsp_dev_sync(), "dev_sync")


def device_memory() -> tuple[int, int]:
    """(free, total) device bytes."""
    f, t = N.C.c_size_t(), N.C.c_size_t()
    check(N.lib.lsp_dev_mem_info(N.C.byref(f), N.C.byref(t)), "mem_info")
    return f.value, t.value


def colamd(c: Circuit) -> np.ndarray:
    q = np.empty(c.n, np.int32)
    check(N.lib.lsp_colamd(c.n, c.nnz, N.ptr(c.row), N.ptr(c.col), N.ptr(q)), "colamd")
    return q


def symbolic_factor(c: Circuit, q=None, tol: float = 1.0):
    """CPU symbolic LU (the pivot sequence and pattern the device reuses).
    Returns dict(pinv, Lp, Li, Lx, Up, Ui, Ux) in the splu layout."""
    q = None if q is None else np.ascontiguousarray(q, np.int32)
    ln, un = N.C.c_long(), N.C.c_long()
    rc = N.lib.lsp_symbolic_factor(c.n, c.nnz, N.ptr(c.row), N.ptr(c.col), N.ptr(c.val),
                                   N.ptr(q), tol, None, None, None, None, 0, None, None, None,
                                   0, N.C.byref(ln), N.C.byref(un))
    if rc >= 0:
        raise RuntimeError(f"symbolic factor: singular column {rc}")
    if rc not in (-1, -3):
        raise RuntimeError(f"symbolic factor failed: {rc}")
    out = dict(pinv=np.empty(c.n, np.int32), Lp=np.empty(c.n + 1, np.int32),
               Li=np.empty(ln.value, np.int32), Lx=np.empty(ln.value),
               Up=np.empty(c.n + 1, np.int32), Ui=np.empty(un.value, np.int32),
               Ux=np.empty(un.value))
    rc = N.lib.lsp_symbolic_factor(c.n, c.nnz, N.ptr(c.row), N.ptr(c.col), N.ptr(c.val),
                                   N.ptr(q), tol, N.ptr(out["pinv"]), N.ptr(out["Lp"]),
                                   N.ptr(out["Li"]), N.ptr(out["Lx"]), ln.value,
                                   N.ptr(out["Up"]), N.ptr(out["Ui"]), N.ptr(out["Ux"]),
                                   un.value, N.C.byref(ln), N.C.byref(un))
    if rc != -1:
        raise RuntimeError(f"symbolic factor failed: {rc}")
    return out


def cpu_lu_ref(c: Circuit, q=None, rhs=None, tol: float = 1.0):
    """The reference's cpu_lu_ref contract: fp64 factor + permuted solve."""
    Ap, Ai, Ax = c.csc(None)
    b = c.rhs if rhs is None else np.ascontiguousarray(rhs, np.float64)
    x = np.empty(c.n)
    ln, un = N.C.c_long(), N.C.c_long()
    qa = None if q is None else np.ascontiguousarray(q, np.int32)
    rc = N.lib.cpu_lu_ref(c.n, N.ptr(Ap), N.ptr(Ai), N.ptr(Ax), N.ptr(qa), N.ptr(b), tol,
                          N.ptr(x), N.C.byref(ln), N.C.byref(un))
    return rc, x, ln.value, un.value


def options(**kw) -> N.Options:
    o = N.Options()
    N.lib.lsp_options_default(N.C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class Batch:
    """Device-resident batch of circuits (one circuit = batch of one).

    Creation runs the host symbolic phase (COLAMD unless q is given, pivot
    sequence, L/U pattern, level sets, SELL layout) and uploads everything;
    refactor()/solve() then run on the GPU."""

    def __init__(self, circuits, upload=True, **opts):
        """upload=False runs only the host symbolic phase; call upload() to
        allocate and upload (estimate_bytes() tells how much it will take)."""
        if isinstance(circuits, Circuit):
            circuits = [circuits]
        self.circuits = list(circuits)
        arr = (N.CircuitIn * len(self.circuits))()
        for i, c in enumerate(self.circuits):
            arr[i].n = c.n
            arr[i].nnz = c.nnz
            arr[i].row = N.ptr(c.row)
            arr[i].col = N.ptr(c.col)
            arr[i].val = N.ptr(c.val)
            arr[i].q = N.ptr(c.q)
        self.opts = options(**opts)
        h = N.C.c_void_p()
        rc = N.lib.lsp_batch_prepare(arr, len(self.circuits), N.C.byref(self.opts), N.C.byref(h))
        check(rc, "lsp_batch_prepare")
        self._h = h
        self.total_n = sum(c.n for c in self.circuits)
        self.total_nnz = sum(c.nnz for c in self.circuits)
        self.uploaded = False
        if upload:
            self.upload()
        self.voff = np.cumsum([0] + [c.n for c in self.circuits])
        self.aoff = np.cumsum([0] + [c.nnz for c in self.circuits])

    def upload(self):
        check(N.lib.lsp_batch_upload(self._h), "lsp_batch_upload")
        self.uploaded = True

    @classmethod
    def adopt(cls, picks, upload=True) -> "Batch":
        """A batch whose circuit i is circuit picks[i][1] of batch picks[i][0],
        moved over with its host structures (no symbolic phase).  Every source
        batch is consumed (closed); circuits not picked are dropped."""
        parts = []
        for b, _ in picks:
            if all(b is not p for p in parts):
                parts.append(b)
        arr = (N.C.c_void_p * len(parts))(*[p._h for p in parts])
        part_of = np.array([next(i for i, p in enumerate(parts) if p is b) for b, _ in picks], np.int32)
        circ_of = np.array([c for _, c in picks], np.int32)
        opts = parts[0].opts
        h = N.C.c_void_p()
        rc = N.lib.lsp_batch_adopt(arr, len(parts), N.ptr(part_of), N.ptr(circ_of), len(picks),
                                   N.C.byref(h))
        for p in parts:
            p._h = None                       # destroyed by the call
        check(rc, "lsp_batch_adopt")
        circuits = [b.circuits[c] for b, c in picks]
        self = cls.__new__(cls)
        self.circuits = circuits
        self.opts = opts
        self._h = h
        self.total_n = sum(c.n for c in circuits)
        self.total_nnz = sum(c.nnz for c in circuits)
        self.uploaded = False
        self.voff = np.cumsum([0] + [c.n for c in circuits])
        self.aoff = np.cumsum([0] + [c.nnz for c in circuits])
        if upload:
            self.upload()
        return self

    def set_active(self, mask=None):
        """Circuits refactor() and solve() work on (None: all)."""
        if mask is None:
            check(N.lib.lsp_batch_set_active(self._h, None), "set_active")
            return
        m = np.ascontiguousarray(np.asarray(mask, bool).astype(np.int32))
        assert m.size == len(self.circuits)
        check(N.lib.lsp_batch_set_active(self._h, N.ptr(m)), "set_active")

    def estimate_bytes(self) -> int:
        v = N.C.c_size_t()
        check(N.lib.lsp_batch_estimate_bytes(self._h, N.C.byref(v)), "estimate_bytes")
        return v.value

    def device_bytes(self) -> int:
        v = N.C.c_size_t()
        check(N.lib.lsp_batch_device_bytes(self._h, N.C.byref(v)), "device_bytes")
        return v.value

    def close(self):
        if getattr(self, "_h", None):
            N.lib.lsp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def info(self, i=0) -> dict:
        inf = N.CircuitInfo()
        check(N.lib.lsp_batch_info(self._h, i, N.C.byref(inf)), "info")
        return {k: getattr(inf, k) for k, _ in N.CircuitInfo._fields_}

    def set_stream(self, stream_handle: int):
        check(N.lib.lsp_batch_set_stream(self._h, N.C.c_void_p(stream_handle)), "set_stream")

    def set_values(self, vals: np.ndarray):
        vals = np.ascontiguousarray(vals, np.float64)
        assert vals.size == self.total_nnz
        check(N.lib.lsp_batch_set_values(self._h, vals.ctypes.data_as(N.C.c_void_p), 0),
              "set_values")

    def set_values_device(self, dptr: int):
        check(N.lib.lsp_batch_set_values(self._h, N.C.c_void_p(dptr), 1), "set_values")

    def refactor(self, prec="fp32") -> dict:
        rep = N.RefactorReport()
        rc = N.lib.lsp_batch_refactor(self._h, prec_code(prec), N.C.byref(rep))
        out = {k: getattr(rep, k) for k, _ in N.RefactorReport._fields_}
        out["rc"] = rc
        if rc != 0:
            out["error"] = N.lib.lsp_batch_last_error(self._h).decode()
        return out

    def trisolve(self, rhs: np.ndarray) -> np.ndarray:
        rhs = np.ascontiguousarray(rhs, np.float64)
        x = np.empty(self.total_n)
        check(N.lib.lsp_batch_trisolve(self._h, rhs.ctypes.data_as(N.C.c_void_p),
                                       x.ctypes.data_as(N.C.c_void_p), 0), "trisolve")
        return x

    def solve(self, rhs: np.ndarray):
        rhs = np.ascontiguousarray(rhs, np.float64)
        x = np.empty(self.total_n)
        rep = N.SolveReport()
        rc = N.lib.lsp_batch_solve(self._h, rhs.ctypes.data_as(N.C.c_void_p),
                                   x.ctypes.data_as(N.C.c_void_p), 0, N.C.byref(rep))
        out = {k: getattr(rep, k) for k, _ in N.SolveReport._fields_}
        out["rc"] = rc
        return x, out

    def pivot_range(self, i=0):
        """(min |u_kk|, max |u_kk|) of circuit i's current factors."""
        lo, hi = N.C.c_double(), N.C.c_double()
        check(N.lib.lsp_batch_pivot_range(self._h, i, N.C.byref(lo), N.C.byref(hi)), "pivot_range")
        return lo.value, hi.value

    def ir_step(self, x: np.ndarray, rhs: np.ndarray):
        """One fused refinement step for every circuit (one launch):
        x + Q U^-1 L^-1 P (rhs - A x) with the current factors; returns the new
        x and ||rhs - A x|| / ||rhs|| of the x passed in, per circuit."""
        rhs = np.ascontiguousarray(rhs, np.float64)
        xn = np.array(x, np.float64, copy=True)
        rel = np.empty(len(self.circuits))
        check(N.lib.lsp_batch_ir_step(self._h, rhs.ctypes.data_as(N.C.c_void_p),
                                      xn.ctypes.data_as(N.C.c_void_p), 0, N.ptr(rel)), "ir_step")
        return xn, rel

    def graph_step(self, prec="fp32"):
        """One Newton step on the resident b / x as a replayed HIP graph
        (refactor + GMRES cycle + residual); returns (refactor, solve) reports."""
        rr = N.RefactorReport()
        sr = N.SolveReport()
        rc = N.lib.lsp_batch_graph_step(self._h, prec_code(prec), N.C.byref(rr), N.C.byref(sr))
        ro = {k: getattr(rr, k) for k, _ in N.RefactorReport._fields_}
        so = {k: getattr(sr, k) for k, _ in N.SolveReport._fields_}
        ro["rc"] = rc
        so["rc"] = rc
        if rc != 0:
            so["error"] = N.lib.lsp_batch_last_error(self._h).decode()
        return ro, so

    def solve_device(self):
        """GMRES-IR on the resident RHS arena (see device_vectors)."""
        b, x = self.device_vectors()
        rep = N.SolveReport()
        rc = N.lib.lsp_batch_solve(self._h, N.C.c_void_p(b), N.C.c_void_p(x), 1,
                                   N.C.byref(rep))
        out = {k: getattr(rep, k) for k, _ in N.SolveReport._fields_}
        out["rc"] = rc
        return out

    def spmv(self, x: np.ndarray, prec="fp64") -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.total_n)
        check(N.lib.lsp_batch_spmv(self._h, x.ctypes.data_as(N.C.c_void_p),
                                   y.ctypes.data_as(N.C.c_void_p), 0, prec_code(prec)), "spmv")
        return y

    def export_factors(self, i=0) -> dict:
        inf = self.info(i)
        n = inf["n"]
        out = dict(pinv=np.empty(n, np.int32), Lp=np.empty(n + 1, np.int32),
                   Li=np.empty(inf["lnnz"], np.int32), Lx=np.empty(inf["lnnz"]),
                   Up=np.empty(n + 1, np.int32), Ui=np.empty(inf["unnz"], np.int32),
                   Ux=np.empty(inf["unnz"]))
        check(N.lib.lsp_batch_export_factors(
            self._h, i, N.ptr(out["pinv"]), N.ptr(out["Lp"]), N.ptr(out["Li"]),
            N.ptr(out["Lx"]), N.ptr(out["Up"]), N.ptr(out["Ui"]), N.ptr(out["Ux"])), "export")
        return out

    def set_rhs(self, src):
        """Copy total_n doubles into the resident RHS arena: src is a numpy
        array or a device address (int)."""
        if isinstance(src, np.ndarray):
            src = np.ascontiguousarray(src, np.float64)
            assert src.size == self.total_n
            check(N.lib.lsp_batch_set_rhs(self._h, src.ctypes.data_as(N.C.c_void_p)), "set_rhs")
        else:
            check(N.lib.lsp_batch_set_rhs(self._h, N.C.c_void_p(int(src))), "set_rhs")

    def get_x(self) -> np.ndarray:
        x = np.empty(self.total_n)
        check(N.lib.lsp_batch_get_x(self._h, x.ctypes.data_as(N.C.c_void_p)), "get_x")
        return x

    def device_vectors(self):
        b, x = N.C.c_void_p(), N.C.c_void_p()
        check(N.lib.lsp_batch_device_vectors(self._h, N.C.byref(b), N.C.byref(x)), "vectors")
        return b.value, x.value

    def kernel_stats(self, reset=False) -> dict:
        st = N.KernelStats()
        check(N.lib.lsp_batch_kernel_stats(self._h, N.C.byref(st), int(reset)), "stats")
        return {k: getattr(st, k) for k, _ in N.KernelStats._fields_}

    def last_times(self):
        a, b, c = N.C.c_double(), N.C.c_double(), N.C.c_double()
        N.lib.lsp_batch_last_times(self._h, N.C.byref(a), N.C.byref(b), N.C.byref(c))
        return a.value, b.value, c.value
