"""Writes the golden fixtures in tests/golden/ from the reference's own C.

Run in the development container, where oracle/_ref/libref.so is compiled
from the reference sources by `make -C oracle ref` (splu.h lu_factor /
lu_solve_perm in fp64 and fp32, cpu_ref_lu.c, level.c, selector.c).  The
fixtures are data only: inputs and the outputs the reference code produced.

    python tests/golden/make_golden.py

Fixtures:
  c0.npz        1020-unknown MNA circuit (seed 0), the shared column order q,
                the reference's fp64 factors (pinv and every L/U array in
                storage order), its fp32-storage factors (splu -DLU_FLOAT),
                x from the reference's cpu_lu_ref, lnnz/unnz and levels.
  tiny.npz      hand-made cases: a source/ground MNA with exact +-1 pivot
                ties, an 8x8 unsymmetric matrix, a structurally singular
                matrix (empty column) and a numerically singular one.
  c1.json       the 316x316 power grid + 50 pads (n = 99906, seed 1): the
                matrix's and q's sha256, and the sha256 of every integer and
                fp64 array the reference's splu, cpu_lu_ref-chain solve and
                level.c produce on it (pinv, Lp, Li, Lx, Up, Ui, Ux, x,
                levels, level histogram), lnnz / unnz / nlevels and a strided
                sample of x.  Hashes keep the fixture small (the arrays are
                tens of MB).
  selector.json the five cases of the reference's bench/test_selector.c:36-87
                with their expected decisions, plus the compiled selector's
                outputs on 200 random feature vectors.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


class MatrixInfo(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_long), ("avg_rowlen", C.c_double),
                ("max_rowlen", C.c_double), ("p50_rowlen", C.c_double),
                ("row_skew", C.c_double), ("kappa_ruiz", C.c_double),
                ("kappa_est", C.c_double), ("dyn_range_log10", C.c_double),
                ("singular_flag", C.c_int)]


class Config(C.Structure):
    _fields_ = [("spmv_format", C.c_int), ("sigma", C.c_int), ("use_ruiz", C.c_int),
                ("use_shift", C.c_int), ("shift", C.c_double), ("factor_prec", C.c_int),
                ("ir_outer", C.c_int), ("ir_inner", C.c_int), ("ir_rtol", C.c_double),
                ("mp_path", C.c_int), ("pre_ruiz_hint", C.c_int), ("fp16_candidate", C.c_int)]


def factor_arrays(F, prefix):
    return {f"{prefix}_{k}": F[k] for k in ("pinv", "Lp", "Li", "Lx", "Up", "Ui", "Ux")}


def c0_fixture():
    import luspice_amd as L
    c = L.mna("c0", seed=0)
    q = L.colamd(c)
    F64 = O.lu_factor(c.n, c.row, c.col, c.val, q, which="ref64")
    F32 = O.lu_factor(c.n, c.row, c.col, c.val, q, which="ref32")
    rc, x, ln, un = O.cpu_lu_ref(c.n, c.row, c.col, c.val, q, c.rhs, which="ref_cpu_lu_ref")
    assert rc == -1
    lv, nl = O.levels(c.n, F64["Up"], F64["Ui"], which="ref")
    out = dict(n=np.int32(c.n), row=c.row, col=c.col, val=c.val, rhs=c.rhs, q=q,
               x=x, lnnz=np.int64(ln), unnz=np.int64(un), level=lv, nlevels=np.int32(nl))
    out.update(factor_arrays(F64, "f64"))
    out.update(factor_arrays(F32, "f32"))
    np.savez_compressed(os.path.join(HERE, "c0.npz"), **out)
    print(f"c0: n={c.n} lnnz={ln} unnz={un} levels={nl}")


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def c1_fixture():
    import luspice_amd as L
    c = L.mna("c1", seed=1)
    q = L.colamd(c)
    F = O.lu_factor(c.n, c.row, c.col, c.val, q, which="ref64")
    assert F["singular"] == -1
    rc, x, ln, un = O.cpu_lu_ref(c.n, c.row, c.col, c.val, q, c.rhs, which="ref64_factor_solve")
    assert rc == -1
    lv, nl = O.levels(c.n, F["Up"], F["Ui"], which="ref")
    hist = O.level_hist(lv, nl, which="ref")
    out = dict(n=c.n, nnz=c.nnz, lnnz=int(ln), unnz=int(un), nlevels=int(nl),
               sha=dict(row=sha(c.row), col=sha(c.col), val=sha(c.val), rhs=sha(c.rhs), q=sha(q),
                        x=sha(x), level=sha(lv), hist=sha(hist),
                        **{k: sha(F[k]) for k in ("pinv", "Lp", "Li", "Lx", "Up", "Ui", "Ux")}),
               x_stride=97, x_sample=[float(v) for v in x[::97]])
    with open(os.path.join(HERE, "c1.json"), "w") as f:
        json.dump(out, f)
    print(f"c1: n={c.n} lnnz={ln} unnz={un} levels={nl}")


def tiny_cases():
    cases = {}
    # 3-node resistor chain, one source at node 0, one at node 2: exact ties
    #   rows 0..2 nodes, rows 3,4 sources (zero diagonal)
    ent = [(0, 0, 2.0), (0, 1, -1.0), (1, 0, -1.0), (1, 1, 2.0), (1, 2, -1.0),
           (2, 1, -1.0), (2, 2, 1.0), (0, 3, 1.0), (3, 0, 1.0), (2, 4, 1.0), (4, 2, 1.0)]
    cases["ties"] = (5, ent, [1.0, 0.0, -1.0, 0.5, 1.0])
    rng = np.random.default_rng(7)
    ent = []
    for i in range(8):
        ent.append((i, i, float(rng.uniform(0.1, 1.0))))
        for j in rng.choice(8, 3, replace=False):
            if j != i:
                ent.append((i, int(j), float(rng.uniform(-2, 2))))
    cases["unsym8"] = (8, ent, list(rng.uniform(-1, 1, 8)))
    # column 2 structurally empty -> singular at the column that holds it
    ent = [(0, 0, 1.0), (1, 1, 2.0), (2, 1, 1.0), (1, 0, 3.0)]
    cases["struct_singular"] = (3, ent, [1.0, 1.0, 1.0])
    # rank deficient: row 2 = row 0 + row 1
    ent = [(0, 0, 1.0), (0, 1, 2.0), (1, 1, 1.0), (1, 2, 1.0), (2, 0, 1.0), (2, 1, 3.0),
           (2, 2, 1.0)]
    cases["num_singular"] = (3, ent, [1.0, 2.0, 3.0])
    out = {}
    for name, (n, ent, b) in cases.items():
        row = np.array([e[0] for e in ent], np.int32)
        col = np.array([e[1] for e in ent], np.int32)
        val = np.array([e[2] for e in ent], np.float64)
        for qname, q in (("nat", None), ("rev", np.arange(n)[::-1].astype(np.int32))):
            key = f"{name}_{qname}"
            F = O.lu_factor(n, row, col, val, q, which="ref64")
            out[f"{key}_n"] = np.int32(n)
            out[f"{key}_row"], out[f"{key}_col"], out[f"{key}_val"] = row, col, val
            out[f"{key}_b"] = np.array(b, np.float64)
            out[f"{key}_q"] = np.arange(n, dtype=np.int32) if q is None else q
            out[f"{key}_singular"] = np.int32(F["singular"])
            if F["singular"] == -1:
                out.update(factor_arrays(F, key))
                rc, x, _, _ = O.cpu_lu_ref(n, row, col, val, q, np.array(b), which="ref_cpu_lu_ref")
                out[f"{key}_x"] = x
    np.savez_compressed(os.path.join(HERE, "tiny.npz"), **out)
    print("tiny:", sorted({k.rsplit('_', 1)[0] for k in out if k.endswith('_singular')}))


def selector_fixture():
    R = O.ref()
    R.dsh_format_select.argtypes = [C.POINTER(MatrixInfo), C.POINTER(Config)]
    R.dsh_mp_escalate.restype = C.c_int
    R.dsh_matrix_analyze.argtypes = [C.c_int, C.c_long, C.POINTER(C.c_int), C.POINTER(MatrixInfo)]
    # bench/test_selector.c:36-87 (features and expected decisions)
    kat = [
        dict(name="U1 bandgap", info=[113, 616, 5.5, 16, 5, 2.9, 7.6e7, 7.6e7, 28.2, 0],
             fmt=0, path=0, pre_ruiz=1, factor_prec=0),
        dict(name="U2 rajat22", info=[39899, 195429, 4.9, 60, 4, 12.2, 2.8e15, 2.8e15, 81.7, 0],
             fmt=0, path=0, pre_ruiz=1),
        dict(name="U3 flash-adc", info=[989, 5798, 5.9, 32, 5, 5.4, 1.5e13, 1.5e13, 46.5, 0],
             path=0),
        dict(name="U4 OTA", info=[42, 210, 5.0, 20, 4, 5.0, 1.5e13, 1.5e13, 27.0, 0], path=0),
        dict(name="U5 rajat16", info=[94294, 432866, 4.6, 50, 4, 10.9, 0, 0, 82.5, 1], path=3),
    ]
    escal = {"0": 1, "1": 2, "2": 3, "3": 3}
    rng = np.random.default_rng(11)
    rand = []
    for _ in range(200):
        info = [int(rng.integers(1, 200000)), int(rng.integers(1, 10 ** 6)),
                float(rng.uniform(1, 10)), float(rng.integers(1, 100)),
                float(rng.integers(1, 10)), float(rng.choice([rng.uniform(0, 12), 5.0, 4.999])),
                float(10 ** rng.uniform(-1, 16)) * float(rng.integers(0, 2)),
                float(10 ** rng.uniform(0, 16)), float(rng.choice([rng.uniform(-1, 90), 1.0, 0.5, 20.0, 20.5])),
                int(rng.integers(0, 2) if rng.uniform() < 0.2 else 0)]
        mi = MatrixInfo(*info)
        cfg = Config()
        R.dsh_format_select(C.byref(mi), C.byref(cfg))
        rand.append(dict(info=info, cfg={k: getattr(cfg, k) for k, _ in Config._fields_}))
    # analyze on random row pointers
    ana = []
    for _ in range(20):
        n = int(rng.integers(1, 300))
        lens = rng.integers(0, 12, n)
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        mi = MatrixInfo()
        R.dsh_matrix_analyze(n, int(rp[-1]), rp.ctypes.data_as(C.POINTER(C.c_int)), C.byref(mi))
        ana.append(dict(rowptr=rp.tolist(), out={k: getattr(mi, k) for k, _ in MatrixInfo._fields_}))
    with open(os.path.join(HERE, "selector.json"), "w") as f:
        json.dump(dict(kat=kat, escalate=escal, random=rand, analyze=ana), f)
    print("selector: 5 KAT + 200 random + 20 analyze")


if __name__ == "__main__":
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libref.so missing: run `make -C oracle ref` first")
    c0_fixture()
    c1_fixture()
    tiny_cases()
    selector_fixture()
