"""Parity on config 1: the 316x316 power grid + 50 pads (n = 99906, seed 1),
the configuration the headline metric is quoted on.

* fp64 factor storage: P and the L/U pattern equal the oracle's exactly, and
  every column of L and U (the T = 1853 dense trailing block included) is
  within 1e-10 of the oracle's column, relative to that column's largest
  entry.  The summation order differs (segment runs vs the oracle's
  depth-first order), so values are compared with a tolerance.
* fp32 factor storage: the factors are an fp32 rounding of an exact
  factorization.  The refactor computes every value in fp64 from the stored
  fp32 values of earlier columns and rounds once when it stores it (the dense
  block is computed in fp64 and rounded once at the end), so
  |P A Q - L U| <= 2 u32 |L| |U| entrywise, u32 = 2^-24, plus fp64
  accumulation error.  Checked through products with a positive vector v:
  |P A Q v - L (U v)| <= (2 u32 + 1e-12) |L| (|U| v).
  The same check with 1e-12 in place of 2 u32 holds the fp64 factors.
* The fp32-factor GMRES-IR solution is within 1e-6 (relative L2) of the
  oracle's fp64 solution (the cpu_lu_ref chain).
"""
import numpy as np
import pytest

import luspice_amd as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24


@pytest.fixture(scope="module")
def c1():
    c = L.mna("c1", seed=1)
    return c, L.colamd(c)


@pytest.fixture(scope="module")
def c1_batch(c1):
    c, q = c1
    print("case: c1 n=99906 batch creation", flush=True)
    B = L.Batch(L.Circuit(c.n, c.row, c.col, c.val, c.rhs, q=q))
    yield B
    B.close()


@pytest.fixture(scope="module")
def c1_oracle(c1):
    c, q = c1
    return O.lu_factor(c.n, c.row, c.col, c.val, q)


def _sorted_columns(F, n):
    """Per column, entries sorted by row (the export sorts; splu keeps its
    depth-first reach order)."""
    out = {}
    for p, i, x in (("Lp", "Li", "Lx"), ("Up", "Ui", "Ux")):
        colid = np.repeat(np.arange(n), np.diff(F[p]))
        o = np.lexsort((F[i], colid))
        out[i] = F[i][o]
        out[x] = F[x][o]
    return out


def _colmax(ptr, v):
    """max |v| over each column's slice (every column is non-empty)."""
    return np.maximum.reduceat(np.abs(v), ptr[:-1])


def _backward_check(c, q, F):
    """max over rows of |P A Q v - L (U v)| / (|L| (|U| v)) for v > 0."""
    import scipy.sparse as sp
    n = c.n
    invq = np.empty(n, np.int64)
    invq[q] = np.arange(n)
    PAQ = sp.csr_matrix((c.val, (F["pinv"][c.row], invq[c.col])), shape=(n, n))
    Lm = sp.csc_matrix((F["Lx"], F["Li"], F["Lp"]), shape=(n, n))
    Um = sp.csc_matrix((F["Ux"], F["Ui"], F["Up"]), shape=(n, n))
    v = np.random.default_rng(3).uniform(0.5, 1.0, n)
    r = np.abs(PAQ @ v - Lm @ (Um @ v))
    bound = abs(Lm) @ (abs(Um) @ v)
    return float(np.max(r / bound))


def test_c1_fp64_factors_match_oracle(c1, c1_batch, c1_oracle):
    c, q = c1
    B, R = c1_batch, c1_oracle
    inf = B.info(0)
    assert inf["tail_T"] > 0, inf        # the dense trailing block is exercised
    print("case: c1 refactor fp64, tail_T", inf["tail_T"], flush=True)
    rep = B.refactor("fp64")
    assert rep["rc"] == 0 and rep["bad_pivots"] == 0, rep
    F = B.export_factors(0)
    n = c.n
    assert np.array_equal(F["pinv"], R["pinv"])
    assert np.array_equal(F["Lp"], R["Lp"]) and np.array_equal(F["Up"], R["Up"])
    S = _sorted_columns(R, n)
    assert np.array_equal(F["Li"], S["Li"]) and np.array_equal(F["Ui"], S["Ui"])
    for x, p in (("Lx", "Lp"), ("Ux", "Up")):
        d = _colmax(F[p], F[x] - S[x])
        m = _colmax(F[p], S[x])
        worst = float(np.max(d / m))
        k = int(np.argmax(d / m))
        print(f"  {x}: worst column {k} rel diff {worst:.3e} (tail starts at {n - inf['tail_T']})",
              flush=True)
        assert worst <= 1e-10, (x, k, worst)
    # the dense block's columns by themselves
    t0 = n - inf["tail_T"]
    for x, p in (("Lx", "Lp"), ("Ux", "Up")):
        a, b = F[p][t0], F[p][n]
        seg = np.repeat(np.arange(t0, n), np.diff(F[p][t0:])) - t0
        diff = np.zeros(inf["tail_T"])
        scale = np.zeros(inf["tail_T"])
        np.maximum.at(diff, seg, np.abs(F[x][a:b] - S[x][a:b]))
        np.maximum.at(scale, seg, np.abs(S[x][a:b]))
        assert np.max(diff / scale) <= 1e-10
    be = _backward_check(c, q, F)
    print(f"  fp64 backward error ratio {be:.3e}", flush=True)
    assert be <= 1e-12


def test_c1_fp32_factors_within_fp32_rounding(c1, c1_batch):
    c, q = c1
    B = c1_batch
    print("case: c1 refactor fp32", flush=True)
    rep = B.refactor("fp32")
    assert rep["rc"] == 0 and rep["bad_pivots"] == 0, rep
    F = B.export_factors(0)
    be = _backward_check(c, q, F)
    print(f"  fp32 backward error ratio {be:.3e} (bound {2 * U32 + 1e-12:.3e})", flush=True)
    assert be <= 2 * U32 + 1e-12
    # and not merely fp64 factors: the fp32 rounding is visible
    assert be > 1e-10
    # every stored value is an fp32 number
    assert np.array_equal(F["Lx"], F["Lx"].astype(np.float32).astype(np.float64))
    assert np.array_equal(F["Ux"], F["Ux"].astype(np.float32).astype(np.float64))


def test_c1_gmres_ir_fp32_factors(c1, c1_batch):
    c, q = c1
    B = c1_batch
    print("case: c1 fp32-factor GMRES-IR", flush=True)
    assert B.refactor("fp32")["rc"] == 0
    x, sr = B.solve(c.rhs)
    rc, xo, _, _ = O.cpu_lu_ref(c.n, c.row, c.col, c.val, q, c.rhs)
    assert rc == -1
    err = O.rel_l2(x, xo)
    print(f"  rel err {err:.3e}", sr, flush=True)
    assert sr["rc"] == 0 and sr["fallback"] == 0 and err <= 1e-6
    assert O.residual(c.n, c.row, c.col, c.val, c.rhs, x) <= 1e-9
