"""s0 text files: writer, Matrix Market converter, and the stage-benchmark
suite runner with JSON aggregation.

The s0 format (docs/06_reproduce.md:105-124 of the reference) is two text
files: the matrix, ``n <n>`` then one ``A <i> <j> <v>`` line per entry
(1-based), and the right-hand side, ``n <n>`` then ``RHS <i> <v>``.  Values
are written with 17 significant digits so that they read back bit for bit.

``mtx_to_s0`` follows the converter embedded in scripts/run_gpu_bench.sh:19-40:
the Matrix Market file is read into CSR (duplicates summed), explicit zeros are
dropped, entries are written row by row in CSR order, and the right-hand side
is b = A * ones.

``run_suite`` follows scripts/run_gpu_bench.sh:42-83: for each matrix it runs
the LU stage CLI (``mp_lu_bench <mtx> <rhs> 0 0 <max_n_st> <reps>``) and the
SpMV CLI (``mp_spmv_bench <mtx> <iters> <groups>``) under a time limit, takes
the last ``JSON:{...}`` line of each, and writes one record per matrix:
``{"matrix": m, "lu": {...} | null, "spmv": {...} | null}`` plus
``lu_status`` / ``spmv_status`` ("ok", "timeout-or-no-output",
"exit <rc>", "parse-fail: ...").  The reference extracts its record with
``grep -o '"matrix":.*'``, which drops the opening brace, so its own
``json.loads`` cannot succeed; here the whole object after ``JSON:`` is
parsed.

Command line::

    python -m luspice_amd.s0tools convert A.mtx [B.mtx ...] --out DIR
    python -m luspice_amd.s0tools run --s0 DIR --names A B --out results.json
"""
import argparse
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "bin")


def write_s0(base, n, row, col, val, rhs=None):
    """Write ``base.mtx`` and ``base.rhs`` (0-based row/col in, 1-based out).
    Without rhs, b = A * ones.  Returns the two paths."""
    row = np.asarray(row, np.int64)
    col = np.asarray(col, np.int64)
    val = np.asarray(val, np.float64)
    if not (len(row) == len(col) == len(val)):
        raise ValueError("row, col and val differ in length")
    if len(row) and (row.min() < 0 or col.min() < 0 or row.max() >= n or col.max() >= n):
        raise ValueError("entry index outside 0..n-1")
    if rhs is None:
        rhs = np.zeros(n)
        np.add.at(rhs, row, val)
    rhs = np.asarray(rhs, np.float64)
    if len(rhs) != n:
        raise ValueError("rhs length differs from n")
    mtx, rhs_path = base + ".mtx", base + ".rhs"
    with open(mtx, "w") as f:
        f.write(f"n {n}\n")
        f.writelines(f"A {i + 1} {j + 1} {v:.17g}\n" for i, j, v in zip(row.tolist(), col.tolist(), val.tolist()))
    with open(rhs_path, "w") as f:
        f.write(f"n {n}\n")
        f.writelines(f"RHS {i + 1} {v:.17g}\n" for i, v in enumerate(rhs.tolist()))
    return mtx, rhs_path


def read_s0(mtx, rhs=None):
    """Parse s0 files into (n, row, col, val, b) with 0-based indices, in file
    order.  Lines that do not start with ``A`` (``RHS``) are skipped, as the
    reference's reader does (src/mp/mp_lu_bench.cu:46-85); a right-hand side
    entry that is missing or reads as exactly 0 becomes 1.0 (:83)."""
    n = None
    rows, cols, vals = [], [], []
    with open(mtx) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "n" and n is None:
                n = int(t[1])
            elif t[0] == "A" and len(t) >= 4:
                rows.append(int(t[1]) - 1)
                cols.append(int(t[2]) - 1)
                vals.append(float(t[3]))
    if n is None:
        raise ValueError(f"{mtx}: no 'n' header")
    b = None
    if rhs is not None:
        b = np.zeros(n)
        with open(rhs) as f:
            for line in f:
                t = line.split()
                if len(t) >= 3 and t[0] == "RHS":
                    i = int(t[1]) - 1
                    if 0 <= i < n:
                        b[i] = float(t[2])
        b[b == 0.0] = 1.0
    return n, np.array(rows, np.int32), np.array(cols, np.int32), np.array(vals), b


def mtx_to_s0(mtx_path, out_dir, name=None):
    """Convert one Matrix Market file to ``out_dir/name.mtx`` and ``.rhs``."""
    import scipy.io
    import scipy.sparse as sp

    A = sp.csr_matrix(scipy.io.mmread(mtx_path))
    if A.shape[0] != A.shape[1]:
        raise ValueError(f"{mtx_path}: not square {A.shape}")
    A.sum_duplicates()
    A.eliminate_zeros()
    A.sort_indices()
    n = A.shape[0]
    row = np.repeat(np.arange(n), np.diff(A.indptr))
    b = A @ np.ones(n)
    name = name or os.path.splitext(os.path.basename(mtx_path))[0]
    os.makedirs(out_dir, exist_ok=True)
    return write_s0(os.path.join(out_dir, name), n, row, A.indices, A.data, b)


def _json_record(text):
    """The object after the last ``JSON:`` marker of a CLI's stdout."""
    lines = [ln for ln in text.splitlines() if ln.startswith("JSON:")]
    if not lines:
        return None, "timeout-or-no-output"
    try:
        return json.loads(lines[-1][len("JSON:"):]), "ok"
    except ValueError as e:
        return None, f"parse-fail: {str(e)[:60]}"


def _run(cmd, timeout):
    try:
        p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                           timeout=timeout)
    except subprocess.TimeoutExpired:
        return None, "timeout-or-no-output"
    rec, status = _json_record(p.stdout)
    if p.returncode != 0 and status == "ok":
        status = f"exit {p.returncode}"
    elif p.returncode != 0 and rec is None:
        status = f"exit {p.returncode}"
    return rec, status


def run_suite(s0_dir, names, out_json, max_n_st=0, reps=3, spmv_iters=100, spmv_groups=4,
              lu_timeout=240, spmv_timeout=120, bin_dir=BIN, st_names=()):
    """Run both stage CLIs over ``names`` and write the aggregated records.
    ``st_names`` get ``max_n_st = 15000`` (the reference runs its
    single-thread kernels only on the two small sky130 matrices)."""
    recs = []
    for m in names:
        mtx, rhs = os.path.join(s0_dir, m + ".mtx"), os.path.join(s0_dir, m + ".rhs")
        mst = 15000 if m in st_names else max_n_st
        lu, lu_status = _run([os.path.join(bin_dir, "mp_lu_bench"), mtx, rhs, "0", "0", str(mst),
                              str(reps)], lu_timeout)
        spmv, spmv_status = _run([os.path.join(bin_dir, "mp_spmv_bench"), mtx, str(spmv_iters),
                                  str(spmv_groups)], spmv_timeout)
        recs.append({"matrix": m, "lu": lu, "lu_status": lu_status, "spmv": spmv,
                     "spmv_status": spmv_status})
        print(f"== {m} done: lu {lu_status}, spmv {spmv_status} ==", flush=True)
    with open(out_json, "w") as f:
        json.dump(recs, f, indent=1)
    print(f"saved {len(recs)} records -> {out_json}", flush=True)
    return recs


def main(argv=None):
    ap = argparse.ArgumentParser(prog="luspice_amd.s0tools")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("convert", help="Matrix Market -> s0")
    c.add_argument("mtx", nargs="+")
    c.add_argument("--out", required=True)
    r = sub.add_parser("run", help="stage benchmarks over s0 files, aggregated to JSON")
    r.add_argument("--s0", required=True)
    r.add_argument("--names", nargs="+", required=True)
    r.add_argument("--out", required=True)
    r.add_argument("--max-n-st", type=int, default=0)
    r.add_argument("--st", nargs="*", default=[], help="matrices that also run the single-thread kernels")
    r.add_argument("--reps", type=int, default=3)
    a = ap.parse_args(argv)
    if a.cmd == "convert":
        for p in a.mtx:
            m, _ = mtx_to_s0(p, a.out)
            print(f"converted {m}")
        return 0
    recs = run_suite(a.s0, a.names, a.out, max_n_st=a.max_n_st, reps=a.reps, st_names=tuple(a.st))
    return 0 if all(x["lu_status"] == "ok" for x in recs) else 1


if __name__ == "__main__":
    sys.exit(main())
