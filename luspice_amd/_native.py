"""ctypes binding of libluspice_amd.so (the C ABI of include/luspice.h and the
reference-compatible headers).

The library is built in-tree by ``__graft_entry__.build()`` (make -C
luspice_amd/csrc).  Importing this module without it raises ImportError: the
package has no CPU fallback for the device path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libluspice_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"luspice_amd native library missing at {LIB_PATH}; build it with "
        "`python -c 'import __graft_entry__ as g; g.build()'` (make -C luspice_amd/csrc)")

lib = C.CDLL(LIB_PATH)

c_int_p = C.POINTER(C.c_int)
c_long_p = C.POINTER(C.c_long)
c_double_p = C.POINTER(C.c_double)


class CircuitIn(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_long), ("row", c_int_p), ("col", c_int_p),
                ("val", c_double_p), ("q", c_int_p)]


class Options(C.Structure):
    _fields_ = [("pivot_tol", C.c_double), ("refactor_ptol", C.c_double),
                ("ir_outer", C.c_int), ("ir_inner", C.c_int),
                ("ir_inner_tol", C.c_double), ("ir_outer_tol", C.c_double),
                ("lds_cap", C.c_int), ("blocks_per_cu", C.c_int), ("chunk_rows", C.c_int),
                ("sell_sigma", C.c_int), ("fallback_fp64", C.c_int),
                ("refactor_blocks", C.c_int), ("trisolve_blocks", C.c_int),
                ("tail_max", C.c_int), ("tail_density", C.c_double), ("ruiz_sweeps", C.c_int),
                ("pivot_scaled", C.c_int), ("poll_backoff", C.c_int),
                ("factor_shift", C.c_double), ("krylov_spmv", C.c_int), ("repivot", C.c_int),
                ("allow_fp16", C.c_int), ("ir_method", C.c_int), ("ir_fused_steps", C.c_int),
                ("refactor_fp64", C.c_int)]


class RefactorReport(C.Structure):
    _fields_ = [("status", C.c_int), ("bad_pivots", C.c_int),
                ("singular_circuits", C.c_int), ("first_bad_circuit", C.c_int),
                ("ms", C.c_double), ("repivoted", C.c_int), ("pivot_mismatch", C.c_int)]


class SolveReport(C.Structure):
    _fields_ = [("status", C.c_int), ("outer", C.c_int), ("inner_total", C.c_int),
                ("converged", C.c_int), ("fallback", C.c_int), ("max_rel_res", C.c_double),
                ("singular_fallback", C.c_int)]


class CircuitInfo(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_long), ("lnnz", C.c_long), ("unnz", C.c_long),
                ("n_updates", C.c_long), ("nlevels", C.c_int), ("nlv_fwd", C.c_int),
                ("nlv_bwd", C.c_int), ("max_col", C.c_int), ("sell_slots", C.c_long),
                ("tail_T", C.c_int), ("tail_updates", C.c_long)]


class KernelStats(C.Structure):
    _fields_ = [("refactor_ms", C.c_double), ("refactor_count", C.c_long),
                ("fwd_ms", C.c_double), ("bwd_ms", C.c_double), ("tri_count", C.c_long), ("graph_steps", C.c_long),
                ("dense_ms", C.c_double)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("lsp_strerror", C.c_char_p, C.c_int)
_sig("lsp_options_default", C.c_int, C.POINTER(Options))
_sig("lsp_batch_create", C.c_int, C.POINTER(CircuitIn), C.c_int, C.POINTER(Options),
     C.POINTER(C.c_void_p))
_sig("lsp_batch_destroy", None, C.c_void_p)
_sig("lsp_batch_prepare", C.c_int, C.POINTER(CircuitIn), C.c_int, C.POINTER(Options),
     C.POINTER(C.c_void_p))
_sig("lsp_batch_estimate_bytes", C.c_int, C.c_void_p, C.POINTER(C.c_size_t))
_sig("lsp_batch_upload", C.c_int, C.c_void_p)
_sig("lsp_batch_device_bytes", C.c_int, C.c_void_p, C.POINTER(C.c_size_t))
_sig("lsp_batch_adopt", C.c_int, C.POINTER(C.c_void_p), C.c_int, c_int_p, c_int_p, C.c_int,
     C.POINTER(C.c_void_p))
_sig("lsp_batch_set_active", C.c_int, C.c_void_p, c_int_p)
_sig("lsp_batch_set_stream", C.c_int, C.c_void_p, C.c_void_p)
_sig("lsp_batch_info", C.c_int, C.c_void_p, C.c_int, C.POINTER(CircuitInfo))
_sig("lsp_batch_sizes", C.c_int, C.c_void_p, c_long_p, c_long_p)
_sig("lsp_batch_set_values", C.c_int, C.c_void_p, C.c_void_p, C.c_int)
_sig("lsp_batch_refactor", C.c_int, C.c_void_p, C.c_int, C.POINTER(RefactorReport))
_sig("lsp_batch_trisolve", C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int)
_sig("lsp_batch_solve", C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
     C.POINTER(SolveReport))
_sig("lsp_batch_spmv", C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int)
_sig("lsp_batch_ir_step", C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, c_double_p)
_sig("lsp_batch_pivot_range", C.c_int, C.c_void_p, C.c_int, c_double_p, c_double_p)
_sig("lsp_ruiz", C.c_int, C.c_int, C.c_long, c_int_p, c_int_p, c_double_p, C.c_int, c_double_p, c_double_p)
_sig("lsp_batch_export_factors", C.c_int, C.c_void_p, C.c_int, c_int_p, c_int_p, c_int_p,
     c_double_p, c_int_p, c_int_p, c_double_p)
_sig("lsp_batch_graph_step", C.c_int, C.c_void_p, C.c_int, C.POINTER(RefactorReport), C.POINTER(SolveReport))
_sig("lsp_batch_device_vectors", C.c_int, C.c_void_p, C.POINTER(C.c_void_p),
     C.POINTER(C.c_void_p))
_sig("lsp_batch_set_rhs", C.c_int, C.c_void_p, C.c_void_p)
_sig("lsp_batch_get_x", C.c_int, C.c_void_p, C.c_void_p)
_sig("lsp_batch_last_times", C.c_int, C.c_void_p, c_double_p, c_double_p, c_double_p)
_sig("lsp_dev_alloc", C.c_int, C.POINTER(C.c_void_p), C.c_size_t)
_sig("lsp_dev_free", C.c_int, C.c_void_p)
_sig("lsp_dev_copy", C.c_int, C.c_void_p, C.c_void_p, C.c_size_t)
_sig("lsp_dev_sync", C.c_int)
_sig("lsp_dev_mem_info", C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t))
_sig("lsp_batch_last_error", C.c_char_p, C.c_void_p)
_sig("lsp_batch_kernel_stats", C.c_int, C.c_void_p, C.POINTER(KernelStats), C.c_int)
_sig("lsp_colamd", C.c_int, C.c_int, C.c_long, c_int_p, c_int_p, c_int_p)
_sig("lsp_symbolic_factor", C.c_int, C.c_int, C.c_long, c_int_p, c_int_p, c_double_p, c_int_p,
     C.c_double, c_int_p, c_int_p, c_int_p, c_double_p, C.c_long, c_int_p, c_int_p, c_double_p,
     C.c_long, c_long_p, c_long_p)
_sig("lsp_mna_generate", C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, c_int_p, c_long_p,
     c_int_p, c_int_p, c_double_p, c_double_p)
_sig("lsp_mna_perturb_values", None, C.c_uint64, C.c_int, C.c_double, C.c_long, c_double_p,
     c_double_p, C.c_int, c_double_p)
_sig("cpu_lu_ref", C.c_int, C.c_int, c_int_p, c_int_p, c_double_p, c_int_p, c_double_p,
     C.c_double, c_double_p, c_long_p, c_long_p)
_sig("dsh_level_compute", C.c_int, C.c_int, c_int_p, c_int_p, c_int_p, c_int_p)
_sig("dsh_level_hist", C.c_int, C.c_int, c_int_p, C.c_int, c_int_p)


def ptr(a: np.ndarray, ctype=None):
    """ctypes pointer to a contiguous numpy array (None passes through)."""
    if a is None:
        return None
    if ctype is None:
        ctype = {np.dtype(np.int32): C.c_int, np.dtype(np.float64): C.c_double,
                 np.dtype(np.float32): C.c_float, np.dtype(np.int64): C.c_long}[a.dtype]
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def check(rc: int, what: str = ""):
    if rc != 0:
        raise RuntimeError(f"{what}: {lib.lsp_strerror(rc).decode()} ({rc})")
    return rc
