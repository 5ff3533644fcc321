"""CPU tests: the oracle and the product's host-side pieces against the
fixtures written by the reference's own C (tests/golden/make_golden.py),
against brute-force dense solves and closed forms, and against invariants.
"""
import json
import os

import numpy as np
import pytest

import luspice_amd as L
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def g0():
    return dict(np.load(os.path.join(GOLD, "c0.npz")))


@pytest.fixture(scope="module")
def tiny():
    return dict(np.load(os.path.join(GOLD, "tiny.npz")))


KEYS = ("pinv", "Lp", "Li", "Lx", "Up", "Ui", "Ux")


def _same(F, g, prefix):
    for k in KEYS:
        assert np.array_equal(F[k], g[f"{prefix}_{k}"]), k


# --------------------------------------------------------------------------
# oracle pinned to the reference
def test_oracle_factors_bitwise_c0(g0):
    F = O.lu_factor(int(g0["n"]), g0["row"], g0["col"], g0["val"], g0["q"], which="oracle")
    _same(F, g0, "f64")


def test_oracle_cpu_lu_ref_bitwise_c0(g0):
    rc, x, ln, un = O.cpu_lu_ref(int(g0["n"]), g0["row"], g0["col"], g0["val"], g0["q"], g0["rhs"])
    assert rc == -1 and ln == g0["lnnz"] and un == g0["unnz"]
    assert np.array_equal(x, g0["x"])


def test_oracle_levels_c0(g0):
    lv, nl = O.levels(int(g0["n"]), g0["f64_Up"], g0["f64_Ui"])
    assert nl == g0["nlevels"] and np.array_equal(lv, g0["level"])


@pytest.mark.parametrize("case", ["ties", "unsym8", "struct_singular", "num_singular"])
@pytest.mark.parametrize("qn", ["nat", "rev"])
def test_oracle_tiny(tiny, case, qn):
    k = f"{case}_{qn}"
    n = int(tiny[f"{k}_n"])
    F = O.lu_factor(n, tiny[f"{k}_row"], tiny[f"{k}_col"], tiny[f"{k}_val"], tiny[f"{k}_q"])
    assert F["singular"] == tiny[f"{k}_singular"]
    if F["singular"] == -1:
        _same(F, tiny, k)
        rc, x, _, _ = O.cpu_lu_ref(n, tiny[f"{k}_row"], tiny[f"{k}_col"], tiny[f"{k}_val"],
                                   tiny[f"{k}_q"], tiny[f"{k}_b"])
        assert np.array_equal(x, tiny[f"{k}_x"])
        # brute force: dense solve of A x = b
        A = np.zeros((n, n))
        np.add.at(A, (tiny[f"{k}_row"], tiny[f"{k}_col"]), tiny[f"{k}_val"])
        np.testing.assert_allclose(A @ x, tiny[f"{k}_b"], atol=1e-12)


def _dense_factors(F, n):
    Ld = np.zeros((n, n))
    Ud = np.zeros((n, n))
    for k in range(n):
        for e in range(F["Lp"][k], F["Lp"][k + 1]):
            Ld[F["Li"][e], k] = F["Lx"][e]
        for e in range(F["Up"][k], F["Up"][k + 1]):
            Ud[F["Ui"][e], k] = F["Ux"][e]
    return Ld, Ud


def test_oracle_closed_form_PAQ_equals_LU(g0):
    """P A(:,q) = L U, with L unit lower, U upper, diagonal positions first/last."""
    n = int(g0["n"])
    F = {k: g0[f"f64_{k}"] for k in KEYS}
    Ld, Ud = _dense_factors(F, n)
    assert np.allclose(np.diag(Ld), 1.0) and np.allclose(np.triu(Ld, 1), 0)
    assert np.allclose(np.tril(Ud, -1), 0)
    for k in range(n):
        assert F["Li"][F["Lp"][k]] == k and F["Ui"][F["Up"][k + 1] - 1] == k
    A = np.zeros((n, n))
    np.add.at(A, (g0["row"], g0["col"]), g0["val"])
    P = np.zeros((n, n))
    P[F["pinv"], np.arange(n)] = 1.0
    PAQ = P @ A[:, g0["q"]]
    assert np.abs(PAQ - Ld @ Ud).max() <= 1e-12 * np.abs(A).max()
    # x = Q U^-1 L^-1 P b
    z = np.linalg.solve(Ud, np.linalg.solve(Ld, P @ g0["rhs"]))
    x = np.empty(n)
    x[g0["q"]] = z
    assert np.linalg.norm(x - g0["x"]) <= 1e-12 * np.linalg.norm(g0["x"])


def test_oracle_trisolve_f32_closed_form(g0):
    n = int(g0["n"])
    F = {k: g0[f"f32_{k}"] for k in KEYS}
    b = g0["rhs"]
    x = O.trisolve_f32(n, F, g0["q"], b)
    Fd = dict(F, Lx=F["Lx"].astype(np.float64), Ux=F["Ux"].astype(np.float64))
    Ld, Ud = _dense_factors(Fd, n)
    P = np.zeros((n, n))
    P[F["pinv"], np.arange(n)] = 1.0
    z = np.linalg.solve(Ud, np.linalg.solve(Ld, P @ b))
    xd = np.empty(n)
    xd[g0["q"]] = z
    assert np.linalg.norm(x - xd) <= 1e-10 * np.linalg.norm(xd)


def test_oracle_gmres_ir_c0(g0):
    """The reference's mixed path (fp32 splu factors + fp64 GMRES-IR), restated,
    reaches the fp64 solution to 1e-6 and a 1e-10 residual."""
    n = int(g0["n"])
    F = {k: g0[f"f32_{k}"] for k in KEYS}
    x, info = O.gmres_ir(n, g0["row"], g0["col"], g0["val"], g0["rhs"], F, g0["q"])
    assert info["rc"] == 0 and info["res"] <= 1e-9
    assert O.rel_l2(x, g0["x"]) <= 1e-6


def test_oracle_gmres_catches_wrong_factor(g0):
    """A factor with one flipped sign still converges in GMRES (it only
    preconditions) but its direct solve misses the oracle by far."""
    n = int(g0["n"])
    F = {k: g0[f"f32_{k}"].copy() for k in KEYS}
    F["Ux"][F["Up"][n // 2]] *= -1.0
    xd = O.trisolve_f32(n, F, g0["q"], g0["rhs"])
    assert O.rel_l2(xd, g0["x"]) > 1e-6


def test_oracle_ruiz_invariants():
    c = L.mna("c0", seed=0)
    v, R, Cs = O.ruiz(c.n, c.row, c.col, c.val, sweeps=30)
    # A_s = diag(R) A diag(C) exactly as applied
    np.testing.assert_allclose(v, R[c.row] * c.val * Cs[c.col], rtol=1e-12)
    rmax = np.zeros(c.n)
    np.maximum.at(rmax, c.row, np.abs(v))
    cmax = np.zeros(c.n)
    np.maximum.at(cmax, c.col, np.abs(v))
    assert np.all(np.abs(rmax - 1) < 1e-3) and np.all(np.abs(cmax - 1) < 1e-3)
    # solving the scaled system recovers x = C x_s
    A = c.dense()
    As = np.zeros_like(A)
    np.add.at(As, (c.row, c.col), v)
    xs = np.linalg.solve(As, R * c.rhs)
    np.testing.assert_allclose(Cs * xs, np.linalg.solve(A, c.rhs), rtol=1e-8, atol=1e-12)


def test_oracle_diag_shift_pins():
    """Factor-only shift (bench/test_lu_gpu_f32.cu:354-364): zero shift is the
    identity; the first (i, i) entry of each row moves away from zero by the
    relative amount s with its sign kept; a later duplicate of that diagonal
    and every off-diagonal entry are untouched; and factors of the shifted
    copy still precondition the unshifted system to the fp64 solution."""
    c = L.mna("c0", seed=0)
    # a duplicate diagonal entry appended for row 7 (rows that already have one)
    has = np.zeros(c.n, bool)
    has[c.row[c.row == c.col]] = True
    i_dup = int(np.flatnonzero(has)[7])
    row = np.append(c.row, i_dup).astype(np.int32)
    col = np.append(c.col, i_dup).astype(np.int32)
    val = np.append(c.val, 0.5)
    assert np.array_equal(O.diag_shift(c.n, row, col, val, 0.0), val)
    s = 1e-3
    out = O.diag_shift(c.n, row, col, val, s)
    changed = np.flatnonzero(out != val)
    # exactly one entry per row that has a diagonal, and it is the first one
    assert len(changed) == int(has.sum())
    assert np.all(row[changed] == col[changed])
    first = {}
    for e in range(len(row)):
        if row[e] == col[e] and row[e] not in first:
            first[int(row[e])] = e
    assert sorted(first.values()) == changed.tolist()
    assert out[-1] == val[-1]                   # the later duplicate
    np.testing.assert_allclose(np.abs(out[changed]), (1 + s) * np.abs(val[changed]), rtol=1e-15)
    assert np.all(np.sign(out[changed]) == np.sign(val[changed]))
    # preconditioning only: fp32 factors of the shifted copy (same q) drive the
    # restated GMRES-IR to the unshifted fp64 solution
    q = L.colamd(c)
    vs = O.diag_shift(c.n, c.row, c.col, c.val, 1e-7)
    F = O.lu_factor(c.n, c.row, c.col, vs, q)
    F32 = dict(F, Lx=F["Lx"].astype(np.float32), Ux=F["Ux"].astype(np.float32))
    x, info = O.gmres_ir(c.n, c.row, c.col, c.val, c.rhs, F32, q)
    rc, xo, _, _ = O.cpu_lu_ref(c.n, c.row, c.col, c.val, q, c.rhs)
    assert rc == -1 and info["rc"] == 0
    assert O.rel_l2(x, xo) <= 1e-6
    # and a shift large enough to matter makes the direct solve miss
    Fb = O.lu_factor(c.n, c.row, c.col, O.diag_shift(c.n, c.row, c.col, c.val, 1e-2), q)
    xd = O.trisolve_f32(c.n, dict(Fb, Lx=Fb["Lx"].astype(np.float32), Ux=Fb["Ux"].astype(np.float32)),
                        q, c.rhs)
    assert O.rel_l2(xd, xo) > 1e-6


def test_oracle_sell_order_rule():
    rng = np.random.default_rng(2)
    lens = rng.integers(0, 9, 200)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    order, slens, offs, tot = O.sell_order(200, rp, 32)
    # descending length; equal lengths in reverse index order (counting sort)
    ref = sorted(range(200), key=lambda i: (-lens[i], -i))
    assert list(order) == ref
    assert tot == sum(32 * max(lens[order[s * 32:(s + 1) * 32]]) for s in range(len(slens)))


# --------------------------------------------------------------------------
# product host pieces pinned to the reference
def test_product_symbolic_bitwise_c0(g0):
    c = L.Circuit(int(g0["n"]), g0["row"], g0["col"], g0["val"], g0["rhs"])
    F = L.symbolic_factor(c, g0["q"])
    _same(F, g0, "f64")
    rc, x, ln, un = L.cpu_lu_ref(c, g0["q"])
    assert rc == -1 and np.array_equal(x, g0["x"]) and ln == g0["lnnz"]


@pytest.mark.parametrize("case", ["ties", "unsym8", "struct_singular", "num_singular"])
def test_product_symbolic_tiny(tiny, case):
    for qn in ("nat", "rev"):
        k = f"{case}_{qn}"
        n = int(tiny[f"{k}_n"])
        c = L.Circuit(n, tiny[f"{k}_row"], tiny[f"{k}_col"], tiny[f"{k}_val"], tiny[f"{k}_b"])
        rc, x, _, _ = L.cpu_lu_ref(c, tiny[f"{k}_q"])
        assert rc == tiny[f"{k}_singular"]
        if rc == -1:
            _same(L.symbolic_factor(c, tiny[f"{k}_q"]), tiny, k)
            assert np.array_equal(x, tiny[f"{k}_x"])


def test_product_levels(g0):
    import ctypes as C
    from luspice_amd import _native as N
    n = int(g0["n"])
    lv = np.empty(n, np.int32)
    nl = C.c_int()
    N.lib.dsh_level_compute(n, N.ptr(g0["f64_Up"]), N.ptr(g0["f64_Ui"]), N.ptr(lv), C.byref(nl))
    assert nl.value == g0["nlevels"] and np.array_equal(lv, g0["level"])
    hist = np.empty(nl.value, np.int32)
    N.lib.dsh_level_hist(n, N.ptr(lv), nl.value, N.ptr(hist))
    assert np.array_equal(hist, np.bincount(lv, minlength=nl.value))


def test_product_selector_kat_and_random():
    import ctypes as C
    from luspice_amd import _native as N
    from tests.golden.make_golden import Config, MatrixInfo
    lib = N.lib
    lib.dsh_format_select.argtypes = [C.POINTER(MatrixInfo), C.POINTER(Config)]
    lib.dsh_mp_escalate.restype = C.c_int
    lib.dsh_mp_name.restype = C.c_char_p
    lib.dsh_fmt_name.restype = C.c_char_p
    lib.dsh_matrix_analyze.argtypes = [C.c_int, C.c_long, C.POINTER(C.c_int), C.POINTER(MatrixInfo)]
    with open(os.path.join(GOLD, "selector.json")) as f:
        G = json.load(f)
    for case in G["kat"]:
        mi, cfg = MatrixInfo(*case["info"]), Config()
        lib.dsh_format_select(C.byref(mi), C.byref(cfg))
        if "fmt" in case:
            assert cfg.spmv_format == case["fmt"], case["name"]
        assert cfg.mp_path == case["path"], case["name"]
        if "pre_ruiz" in case:
            assert cfg.pre_ruiz_hint == case["pre_ruiz"]
        if "factor_prec" in case:
            assert cfg.factor_prec == case["factor_prec"]
    for cur, nxt in G["escalate"].items():
        assert lib.dsh_mp_escalate(int(cur)) == nxt
    assert lib.dsh_mp_name(0) == b"P-fast" and lib.dsh_fmt_name(1) == b"SELL-C-sigma32"
    for r in G["random"]:
        mi, cfg = MatrixInfo(*r["info"]), Config()
        lib.dsh_format_select(C.byref(mi), C.byref(cfg))
        got = {k: getattr(cfg, k) for k, _ in Config._fields_}
        assert got == r["cfg"]
    for a in G["analyze"]:
        rp = np.array(a["rowptr"], np.int32)
        mi = MatrixInfo()
        lib.dsh_matrix_analyze(len(rp) - 1, int(rp[-1]), N.ptr(rp), C.byref(mi))
        got = {k: getattr(mi, k) for k, _ in MatrixInfo._fields_}
        assert got == pytest.approx(a["out"])


# --------------------------------------------------------------------------
# invariants of the unpinned pieces (COLAMD, generator)
def test_colamd_invariants():
    c = L.mna("c0", seed=0)
    q = L.colamd(c)
    assert sorted(q) == list(range(c.n))
    assert np.array_equal(q, L.colamd(c))                     # deterministic
    Fq = L.symbolic_factor(c, q)
    Fn = L.symbolic_factor(c, None)
    assert Fq["Lp"][-1] + Fq["Up"][-1] < 0.5 * (Fn["Lp"][-1] + Fn["Up"][-1])
    # a column with more than 16 entries is dense for all-zero knobs: last
    n = 40
    row, col, val = [], [], []
    for i in range(n):
        row.append(i); col.append(i); val.append(4.0)
        if i + 1 < n:
            row += [i, i + 1]; col += [i + 1, i]; val += [-1.0, -1.0]
    for i in range(20):
        row.append(i); col.append(7); val.append(0.5)
    d = L.Circuit(n, row, col, val)
    qd = L.colamd(d)
    assert qd[-1] == 7 and sorted(qd) == list(range(n))


def test_mna_generator_invariants():
    c = L.mna("c0", seed=0)
    assert c.n == 1020
    A = c.dense()
    assert np.array_equal(A != 0, (A != 0).T)                   # symmetric pattern
    src_rows = range(1000, 1020)
    assert all(A[r, r] == 0 for r in src_rows)                  # zero-diagonal rows
    assert all(np.count_nonzero(A[r]) == 1 for r in src_rows)
    c2 = L.mna("c0", seed=0)
    assert np.array_equal(c.val, c2.val) and np.array_equal(c.rhs, c2.rhs)
    assert not np.array_equal(L.mna("c0", seed=1).val, c.val)
    # node rows are diagonally dominant (capacitors to ground)
    nodes = A[:1000, :1000]
    off = np.abs(nodes).sum(1) - np.abs(np.diag(nodes))
    assert np.all(np.diag(nodes) > off)
    s = L.mna("sized", n=777, seed=3)
    assert s.n == 777


def test_perturb_is_seeded_lognormal():
    c = L.mna("sized", n=300, seed=1)
    v1, b1 = L.perturb(5, 2, c.val, c.n)
    v2, b2 = L.perturb(5, 2, c.val, c.n)
    assert np.array_equal(v1, v2) and np.array_equal(b1, b2)
    r = np.log(v1 / c.val)
    assert abs(r.mean()) < 0.02 and abs(r.std() - 0.1) < 0.02
    assert np.all(np.abs(b1) <= 1)


# --------------------------------------------------------------------------
# config 1 (n = 99906): integer structure and fp64 values pinned by the
# hashes the reference's own splu / level.c wrote (tests/golden/c1.json)
@pytest.fixture(scope="module")
def g1():
    with open(os.path.join(GOLD, "c1.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def c1_case(g1):
    c = L.mna("c1", seed=1)
    q = L.colamd(c)
    return c, q


def _sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_c1_inputs_match_fixture(g1, c1_case):
    """Generator and COLAMD are deterministic: the matrix, RHS and q the
    fixture was written from are reproduced byte for byte."""
    c, q = c1_case
    assert c.n == g1["n"] and c.nnz == g1["nnz"]
    for k, a in (("row", c.row), ("col", c.col), ("val", c.val), ("rhs", c.rhs), ("q", q)):
        assert _sha(a) == g1["sha"][k], k


def test_product_symbolic_bitwise_c1(g1, c1_case):
    c, q = c1_case
    F = L.symbolic_factor(c, q)
    assert len(F["Li"]) == g1["lnnz"] and len(F["Ui"]) == g1["unnz"]
    for k in KEYS:
        assert _sha(F[k]) == g1["sha"][k], k
    import ctypes as C
    from luspice_amd import _native as N
    lv = np.empty(c.n, np.int32)
    nl = C.c_int()
    N.lib.dsh_level_compute(c.n, N.ptr(F["Up"]), N.ptr(F["Ui"]), N.ptr(lv), C.byref(nl))
    assert nl.value == g1["nlevels"] and _sha(lv) == g1["sha"]["level"]
    hist = np.empty(nl.value, np.int32)
    N.lib.dsh_level_hist(c.n, N.ptr(lv), nl.value, N.ptr(hist))
    assert _sha(hist) == g1["sha"]["hist"]


def test_product_cpu_lu_ref_bitwise_c1(g1, c1_case):
    c, q = c1_case
    rc, x, ln, un = L.cpu_lu_ref(c, q)
    assert rc == -1 and ln == g1["lnnz"] and un == g1["unnz"]
    assert _sha(x) == g1["sha"]["x"]


def test_oracle_bitwise_c1(g1, c1_case):
    c, q = c1_case
    F = O.lu_factor(c.n, c.row, c.col, c.val, q)
    for k in KEYS:
        assert _sha(F[k]) == g1["sha"][k], k
    lv, nl = O.levels(c.n, F["Up"], F["Ui"])
    assert nl == g1["nlevels"] and _sha(lv) == g1["sha"]["level"]
    hist = O.level_hist(lv, nl)
    assert _sha(hist) == g1["sha"]["hist"]
    assert np.array_equal(hist, np.bincount(lv, minlength=nl))
    rc, x, _, _ = O.cpu_lu_ref(c.n, c.row, c.col, c.val, q, c.rhs)
    assert rc == -1 and _sha(x) == g1["sha"]["x"]
    assert np.array_equal(x[::g1["x_stride"]], np.array(g1["x_sample"]))


# --------------------------------------------------------------------------
# the SpMV, residual and level-histogram checkers, pinned by what the
# mathematics fixes
def _rand_coo(rng, n, per_row, lo=-9, hi=10, dup=True):
    row = np.repeat(np.arange(n), per_row).astype(np.int32)
    col = rng.integers(0, n, row.size).astype(np.int32)
    val = rng.integers(lo, hi, row.size).astype(np.float64)
    if dup:                                     # a repeated (i, j) entry is summed
        row = np.append(row, row[:5]).astype(np.int32)
        col = np.append(col, col[:5]).astype(np.int32)
        val = np.append(val, [3.0, -2.0, 7.0, 1.0, -5.0])
    return row, col, val


def _csr(n, row, col, val):
    order = np.lexsort((np.arange(len(row)), row))
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, row + 1, 1)
    return np.cumsum(rp).astype(np.int32), col[order], val[order]


def test_oracle_csr_spmv_exact_on_integers():
    """Small integers keep every product and partial sum exact, so y must equal
    the dense product bit for bit: a wrong column index, a transposed operand or
    a dropped duplicate changes an exact integer."""
    rng = np.random.default_rng(21)
    n = 300
    row, col, val = _rand_coo(rng, n, 6)
    A = np.zeros((n, n))
    np.add.at(A, (row, col), val)
    x = rng.integers(-50, 51, n).astype(np.float64)
    rp, ci, vv = _csr(n, row, col, val)
    y = O.csr_spmv(n, rp, ci, vv, x)
    assert np.array_equal(y, A @ x)
    assert not np.array_equal(y, A.T @ x)
    # an empty first row gives 0; the other rows shift down by one
    y2 = O.csr_spmv(n, np.concatenate([[0, 0], rp[1:n]]).astype(np.int32), ci, vv, x)
    assert y2[0] == 0.0 and np.array_equal(y2[1:], y[:-1])


def test_oracle_csr_spmv_real_values():
    rng = np.random.default_rng(22)
    n = 500
    row, col, _ = _rand_coo(rng, n, 5)
    val = rng.uniform(-1, 1, row.size)
    A = np.zeros((n, n))
    np.add.at(A, (row, col), val)
    x = rng.uniform(-1, 1, n)
    rp, ci, vv = _csr(n, row, col, val)
    y = O.csr_spmv(n, rp, ci, vv, x)
    np.testing.assert_allclose(y, A @ x, rtol=1e-13, atol=1e-14)


def test_oracle_residual_closed_form():
    """||A x - b|| / ||b|| with b = A x + d e_i (exact in integers) is
    |d| / ||b||; b = A x gives exactly 0; the COO entries are 1-based."""
    rng = np.random.default_rng(23)
    n = 200
    row, col, val = _rand_coo(rng, n, 4)
    A = np.zeros((n, n))
    np.add.at(A, (row, col), val)
    x = rng.integers(-20, 21, n).astype(np.float64)
    b = A @ x
    assert O.residual(n, row, col, val, b, x) == 0.0
    for i, d in ((0, 3.0), (137, -0.5), (n - 1, 2.0 ** -10)):   # dyadic: b + d e_i is exact
        bb = b.copy()
        bb[i] += d
        want = abs(d) / np.linalg.norm(bb)
        assert abs(O.residual(n, row, col, val, bb, x) - want) <= 1e-15 * want + 1e-300
    # the residual of a wrong x scales with the error
    xe = x.copy()
    xe[5] += 1.0
    want = np.linalg.norm(A[:, 5]) / np.linalg.norm(b)
    assert abs(O.residual(n, row, col, val, b, xe) - want) <= 1e-14 * want


def test_oracle_rel_l2_closed_form():
    a = np.array([3.0, 4.0, 0.0])
    b = np.array([0.0, 0.0, 5.0])
    assert O.rel_l2(a, b) == np.sqrt(50.0) / 5.0
    assert O.rel_l2(b, b) == 0.0


def test_oracle_level_hist_matches_compiled_reference_and_bincount():
    rng = np.random.default_rng(24)
    for nl in (1, 7, 40):
        lv = rng.integers(0, nl, 1000).astype(np.int32)
        h = O.level_hist(lv, nl)
        assert np.array_equal(h, np.bincount(lv, minlength=nl))
        if O.ref_available():
            assert np.array_equal(h, O.level_hist(lv, nl, which="ref"))
