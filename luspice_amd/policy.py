"""Precision escalation chain for one circuit: P-fast -> P-safe -> P-hard ->
P-sing.

The reference defines the chain (docs/04_adaptive_policy.md:7-30) and the
transition function (src/selector.c:55-90: ``dsh_mp_initial`` starts at
P-fast unless the matrix is flagged singular, ``dsh_mp_escalate`` moves one
step and stops at P-sing), but nothing in the reference calls them.  Here
each path is a configuration of the device-resident batch, tried in order
until the relative residual ||b - A x|| / ||b|| on the fp64 A meets ``tol``
(the reference's pass test, docs/04_adaptive_policy.md:22-24):

* P-fast: fp32 factor storage, fp64 pivot test and fp64 GMRES-IR;
* P-safe: fp64 factors, pivot order chosen on the Ruiz-equilibrated matrix
  (``pivot_scaled``), factor-only diagonal shift ``shift * |a_ii|``
  (default 1e-7, src/selector.c:121-122), GMRES-IR;
* P-hard: fp64 factors, Ruiz pivot order, no shift, a GMRES twice as long
  per restart with three times the restarts;
* P-sing: nothing left to try.  :class:`Unsolved` is raised with every
  attempt's report: the reference's "detect, then report" end of the chain.

Every attempt reuses nothing from the previous one except the circuit; the
IR parameters come from the selector's configuration (src/selector.c:124-126:
3 restarts, 50 inner iterations, 1e-9), and the in-batch fp64 fallback is
off, because the chain itself is the fallback.

Two guards of the reference's precision budget sit on the chain:

* pivot guard (docs/02_precision_budget.md:82-83): after the P-fast refactor,
  min |u_kk| < 1e-9 max |A| marks the fp32 factors unreliable; the solve is
  skipped and the chain moves on to fp64 factors;
* Ruiz rule (docs/03_gmres_ir.md:60-67): Ruiz is used only when it reduces the
  matrix's scale spread by more than 1e2 (row/column imbalance); below that
  (an intrinsically near-singular matrix) P-safe and P-hard pivot on the
  unscaled matrix.  The reduction is measured as the ratio of the entries'
  dynamic range max |a| / min |a| before and after 8 Ruiz sweeps: a diagonal
  scaling D1 A D2 that Ruiz undoes widens both the dynamic range and the
  condition number by the same factors, and the range needs no factorisation.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

P_FAST, P_SAFE, P_HARD, P_SING = 0, 1, 2, 3
PIVOT_GUARD = 1e-9          # min |pivot| < PIVOT_GUARD * max |A|: fp32 factors unreliable
RUIZ_MIN_REDUCTION = 1e2    # Ruiz only when it shrinks the scale spread more than this


class MatrixInfo(C.Structure):          # selector.h dsh_matrix_info
    _fields_ = [("n", C.c_int), ("nnz", C.c_long), ("avg_rowlen", C.c_double),
                ("max_rowlen", C.c_double), ("p50_rowlen", C.c_double), ("row_skew", C.c_double),
                ("kappa_ruiz", C.c_double), ("kappa_est", C.c_double),
                ("dyn_range_log10", C.c_double), ("singular_flag", C.c_int)]


class DshConfig(C.Structure):           # selector.h dsh_config
    _fields_ = [("spmv_format", C.c_int), ("sigma", C.c_int), ("use_ruiz", C.c_int),
                ("use_shift", C.c_int), ("shift", C.c_double), ("factor_prec", C.c_int),
                ("ir_outer", C.c_int), ("ir_inner", C.c_int), ("ir_rtol", C.c_double),
                ("mp_path", C.c_int), ("pre_ruiz_hint", C.c_int), ("fp16_candidate", C.c_int)]


# private function objects (lib["name"] is not cached), so that other
# bindings of the same symbols keep their own signatures
_analyze = N.lib["dsh_matrix_analyze"]
_analyze.argtypes = [C.c_int, C.c_long, C.POINTER(C.c_int), C.POINTER(MatrixInfo)]
_analyze.restype = None
_format_select = N.lib["dsh_format_select"]
_format_select.argtypes = [C.POINTER(MatrixInfo), C.POINTER(DshConfig)]
_format_select.restype = None
_escalate = N.lib["dsh_mp_escalate"]
_escalate.argtypes = [C.c_int]
_escalate.restype = C.c_int
_mp_name = N.lib["dsh_mp_name"]
_mp_name.argtypes = [C.c_int]
_mp_name.restype = C.c_char_p


class Unsolved(RuntimeError):
    """No path of the chain met the tolerance; ``attempts`` holds their reports."""

    def __init__(self, msg, attempts):
        super().__init__(msg)
        self.attempts = attempts


def path_name(p: int) -> str:
    return _mp_name(int(p)).decode()


def escalate(p: int) -> int:
    return int(_escalate(int(p)))


def select(circuit, singular_flag: int = 0) -> DshConfig:
    """The selector's configuration for the circuit's row-length profile."""
    rp = np.zeros(circuit.n + 1, np.int32)
    np.add.at(rp, np.asarray(circuit.row) + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    info = MatrixInfo()
    _analyze(circuit.n, int(circuit.nnz), rp.ctypes.data_as(C.POINTER(C.c_int)), C.byref(info))
    info.singular_flag = int(singular_flag)
    cfg = DshConfig()
    _format_select(C.byref(info), C.byref(cfg))
    return cfg


def _dyn_range(v: np.ndarray) -> float:
    a = np.abs(v)
    a = a[a > 0]
    return float(a.max() / a.min()) if a.size else 1.0


def ruiz_reduction(circuit, sweeps: int = 8) -> float:
    """How much Ruiz equilibration shrinks the scale spread of A: dynamic
    range max |a| / min |a| of A over that of R A C (8 sweeps, the
    reference's own sweep count in results/raw/fp16/cpu_fp16_analysis.py)."""
    row = np.ascontiguousarray(circuit.row, np.int32)
    col = np.ascontiguousarray(circuit.col, np.int32)
    val = np.ascontiguousarray(circuit.val, np.float64)
    R = np.empty(circuit.n)
    Cs = np.empty(circuit.n)
    N.check(N.lib.lsp_ruiz(circuit.n, int(circuit.nnz), N.ptr(row), N.ptr(col), N.ptr(val), int(sweeps),
                           N.ptr(R), N.ptr(Cs)), "lsp_ruiz")
    return _dyn_range(val) / _dyn_range(val * R[row] * Cs[col])


def path_options(p: int, cfg: DshConfig, shift: float | None = None,
                 use_ruiz: bool = True) -> tuple[str, dict]:
    """(factor precision, batch options) of one path of the chain; use_ruiz
    is the Ruiz rule's verdict (pivot order of P-safe / P-hard on R A C)."""
    ir = dict(ir_outer=cfg.ir_outer, ir_inner=cfg.ir_inner, ir_inner_tol=cfg.ir_rtol,
              ir_outer_tol=cfg.ir_rtol, fallback_fp64=0)
    s = cfg.shift if shift is None else shift
    rz = 1 if use_ruiz else 0
    if p == P_FAST:
        return "fp32", ir
    if p == P_SAFE:
        return "fp64", dict(ir, pivot_scaled=rz, factor_shift=s if cfg.use_shift else 0.0)
    if p == P_HARD:
        return "fp64", dict(ir, pivot_scaled=rz, ir_outer=3 * cfg.ir_outer, ir_inner=2 * cfg.ir_inner)
    raise ValueError(f"no solve configuration for {path_name(p)}")


def solve_adaptive(circuit, rhs=None, tol: float = 1e-6, shift: float | None = None,
                   start: int | None = None, batch_factory=None, **extra):
    """Solve one circuit, escalating until ||b - A x|| / ||b|| <= tol.

    Returns (x, report); report["path"] names the path that passed and
    report["attempts"] lists every attempt.  Raises :class:`Unsolved` at P-sing.
    ``batch_factory(circuit, **options)`` builds the batch (default
    :class:`luspice_amd.Batch`); ``extra`` options go to every attempt."""
    if batch_factory is None:
        from . import Batch as batch_factory
    b = circuit.rhs if rhs is None else np.asarray(rhs, np.float64)
    cfg = select(circuit)
    p = cfg.mp_path if start is None else int(start)
    reduction = ruiz_reduction(circuit)
    use_ruiz = reduction > RUIZ_MIN_REDUCTION
    amax = float(np.abs(circuit.val).max()) if circuit.nnz else 0.0
    attempts = []
    while p != P_SING:
        prec, opts = path_options(p, cfg, shift, use_ruiz)
        opts.update(extra)
        rec = dict(path=path_name(p), prec=prec, factor_shift=opts.get("factor_shift", 0.0),
                   pivot_scaled=opts.get("pivot_scaled", 0), ruiz_reduction=reduction)
        B = None
        try:
            B = batch_factory(circuit, **opts)
            rr = B.refactor(prec)
            rec.update(refactor_rc=rr["rc"], bad_pivots=rr.get("bad_pivots", 0))
            if rr["rc"] == 0 and prec == "fp32":
                lo, _ = B.pivot_range(0)
                rec["min_pivot"] = lo
                if not lo >= PIVOT_GUARD * amax:          # NaN fails it too
                    rec["pivot_guard"] = True
                    attempts.append(rec)
                    p = escalate(p)
                    continue
            if rr["rc"] == 0:
                x, sr = B.solve(b)
                rec.update(solve_rc=sr["rc"], rel_res=float(sr["max_rel_res"]),
                           inner=int(sr["inner_total"]), outer=int(sr["outer"]))
                if sr["rc"] == 0 and np.isfinite(sr["max_rel_res"]) and sr["max_rel_res"] <= tol:
                    attempts.append(rec)
                    return x, dict(path=rec["path"], attempts=attempts, config=cfg)
        except RuntimeError as e:             # the symbolic phase found the matrix singular
            rec.update(error=str(e))
        finally:
            if B is not None and hasattr(B, "close"):
                B.close()
        attempts.append(rec)
        p = escalate(p)
    raise Unsolved(f"no path met ||r||/||b|| <= {tol:g}: " +
                   ", ".join(f"{a['path']} {a.get('rel_res', a.get('error', 'refactor failed'))}"
                             for a in attempts), attempts)
