"""The precision escalation chain (luspice_amd.policy).

CPU: the transition order is the compiled selector's (src/selector.c:70-90,
whose five KAT cases are pinned in test_oracle_golden.py), each path's batch
configuration, and the chain's control flow driven by a stand-in batch whose
residuals are scripted.  GPU: the chain on config-0 and config-2 circuits,
from P-fast and forced to start at P-safe and P-hard, against the fp64 oracle.
"""
import numpy as np
import pytest

import luspice_amd as L
from luspice_amd import policy as P


def test_transitions_follow_the_selector():
    assert [P.escalate(p) for p in (P.P_FAST, P.P_SAFE, P.P_HARD, P.P_SING)] == \
        [P.P_SAFE, P.P_HARD, P.P_SING, P.P_SING]
    assert [P.path_name(p) for p in range(4)] == ["P-fast", "P-safe", "P-hard", "P-sing"]
    c = L.mna("c0", seed=0)
    cfg = P.select(c)
    assert cfg.mp_path == P.P_FAST and cfg.use_shift == 1 and cfg.shift == 1e-7
    assert (cfg.ir_outer, cfg.ir_inner, cfg.ir_rtol) == (3, 50, 1e-9)
    assert P.select(c, singular_flag=1).mp_path == P.P_SING


def test_path_options():
    cfg = P.select(L.mna("c0", seed=0))
    prec, o = P.path_options(P.P_FAST, cfg)
    assert prec == "fp32" and o.get("factor_shift", 0.0) == 0.0 and o["fallback_fp64"] == 0
    prec, o = P.path_options(P.P_SAFE, cfg)
    assert prec == "fp64" and o["factor_shift"] == 1e-7 and o["pivot_scaled"] == 1
    prec, o = P.path_options(P.P_HARD, cfg)
    assert prec == "fp64" and o.get("factor_shift", 0.0) == 0.0 and o["ir_outer"] == 9 and o["ir_inner"] == 100
    with pytest.raises(ValueError):
        P.path_options(P.P_SING, cfg)


class _Stub:
    """A batch whose refactor / solve results are scripted per precision and
    options; it records every configuration it was built with."""
    built = []

    def __init__(self, script):
        self.script = script

    def __call__(self, circuit, **opts):
        key = "hard" if opts["ir_outer"] > 3 else "safe" if opts.get("factor_shift") else "fast"
        _Stub.built.append((key, opts))
        s = self.script[key]
        if s == "singular":
            raise RuntimeError("symbolic factor: singular column 3")
        return _StubBatch(circuit.n, s)


class _StubBatch:
    def __init__(self, n, s):
        self.n, self.s = n, s
        self.solved = False

    def pivot_range(self, i=0):
        return (1e-30, 1.0) if self.s == "tiny-pivot" else (0.5, 2.0)

    def refactor(self, prec):
        return {"rc": 0 if self.s != "refactor-fail" else 3, "bad_pivots": 0}

    def solve(self, b):
        res = self.s if isinstance(self.s, float) else 1.0
        return np.full(self.n, res), {"rc": 0, "max_rel_res": res, "inner_total": 4, "outer": 1}

    def close(self):
        pass


@pytest.mark.parametrize("script,path,n_att", [
    ({"fast": 1e-11, "safe": 1e-11, "hard": 1e-11}, "P-fast", 1),
    ({"fast": 3e-4, "safe": 2e-9, "hard": 1e-11}, "P-safe", 2),
    ({"fast": "singular", "safe": "refactor-fail", "hard": 5e-7}, "P-hard", 3),
    ({"fast": float("nan"), "safe": 1e-6, "hard": 1e-11}, "P-safe", 2),
])
def test_chain_control_flow(script, path, n_att):
    _Stub.built = []
    c = L.mna("c0", seed=0)
    x, rep = P.solve_adaptive(c, batch_factory=_Stub(script))
    assert rep["path"] == path and len(rep["attempts"]) == n_att
    assert [k for k, _ in _Stub.built] == ["fast", "safe", "hard"][:n_att]
    assert rep["attempts"][-1]["rel_res"] <= 1e-6


def test_chain_ends_at_p_sing():
    c = L.mna("c0", seed=0)
    with pytest.raises(P.Unsolved) as e:
        P.solve_adaptive(c, batch_factory=_Stub({"fast": 1e-2, "safe": "singular", "hard": 2e-6}))
    att = e.value.attempts
    assert [a["path"] for a in att] == ["P-fast", "P-safe", "P-hard"]
    assert "singular" in att[1]["error"] and att[2]["rel_res"] == 2e-6


@pytest.mark.gpu
def test_chain_on_device():
    from oracle import oracle as O
    c = L.mna("c0", seed=0)
    q = L.colamd(c)
    circ = L.Circuit(c.n, c.row, c.col, c.val, c.rhs, q=q)
    rc, xo, _, _ = O.cpu_lu_ref(c.n, c.row, c.col, c.val, q, c.rhs)
    for start, name in ((None, "P-fast"), (P.P_SAFE, "P-safe"), (P.P_HARD, "P-hard")):
        print(f"case: c0 escalation chain from {name}", flush=True)
        x, rep = P.solve_adaptive(circ, start=start)
        print(rep["attempts"], flush=True)
        assert rep["path"] == name
        assert O.rel_l2(x, xo) <= 1e-6
    for seed in (2001, 2003, 2007):
        print(f"case: c2 n=1500 seed {seed} escalation chain", flush=True)
        c2 = L.mna("c2", n=1500, seed=seed)
        circ = L.Circuit(c2.n, c2.row, c2.col, c2.val, c2.rhs, q=L.colamd(c2))
        x, rep = P.solve_adaptive(circ)
        print(rep["path"], rep["attempts"], flush=True)
        r = O.residual(c2.n, c2.row, c2.col, c2.val, c2.rhs, x)
        assert r <= 1e-6


def test_ruiz_reduction_closed_forms():
    """A matrix whose entries all have |a| = 1 has nothing to equilibrate
    (reduction 1); a diagonal matrix is equilibrated by the sweeps to within
    |d|^(1/2^16) of 1, so the reduction is its dynamic range; a random
    two-sided scaling of c0 is undone far beyond the 1e2 threshold, while c0
    itself and a c2 circuit (intrinsic spread) stay below it."""
    rng = np.random.default_rng(1)
    n = 200
    row = np.r_[np.arange(n), np.arange(n - 1), np.arange(1, n)].astype(np.int32)
    col = np.r_[np.arange(n), np.arange(1, n), np.arange(n - 1)].astype(np.int32)
    v = np.where(rng.random(row.size) < 0.5, -1.0, 1.0)
    assert P.ruiz_reduction(L.Circuit(n, row, col, v, np.ones(n))) == 1.0
    d = 10.0 ** rng.uniform(-4, 4, n)
    dr = d.max() / d.min()
    red = P.ruiz_reduction(L.Circuit(n, np.arange(n, dtype=np.int32), np.arange(n, dtype=np.int32), d,
                                     np.ones(n)))
    assert abs(red / dr - 1.0) <= 1e-3
    c = L.mna("c0", seed=0)
    assert P.ruiz_reduction(c) < P.RUIZ_MIN_REDUCTION
    assert P.ruiz_reduction(L.mna("c2", n=3000, seed=2)) < P.RUIZ_MIN_REDUCTION
    d1, d2 = 10.0 ** rng.uniform(-3, 3, c.n), 10.0 ** rng.uniform(-2, 2, c.n)
    cs = L.Circuit(c.n, c.row, c.col, c.val * d1[c.row] * d2[c.col], c.rhs)
    assert P.ruiz_reduction(cs) > 1e6


def test_ruiz_rule_sets_the_pivot_order_of_the_fp64_paths():
    cfg = P.select(L.mna("c0", seed=0))
    for p in (P.P_SAFE, P.P_HARD):
        assert P.path_options(p, cfg, use_ruiz=True)[1]["pivot_scaled"] == 1
        assert P.path_options(p, cfg, use_ruiz=False)[1]["pivot_scaled"] == 0
    _Stub.built = []
    c = L.mna("c0", seed=0)                       # reduction < 1e2: no Ruiz
    x, rep = P.solve_adaptive(c, batch_factory=_Stub({"fast": 1e-3, "safe": 1e-11, "hard": 1e-11}))
    assert rep["path"] == "P-safe" and _Stub.built[1][1]["pivot_scaled"] == 0
    rng = np.random.default_rng(4)
    d1 = 10.0 ** rng.uniform(-3, 3, c.n)
    cs = L.Circuit(c.n, c.row, c.col, c.val * d1[c.row], c.rhs)
    _Stub.built = []
    x, rep = P.solve_adaptive(cs, batch_factory=_Stub({"fast": 1e-3, "safe": 1e-11, "hard": 1e-11}))
    assert rep["path"] == "P-safe" and _Stub.built[1][1]["pivot_scaled"] == 1
    assert rep["attempts"][0]["ruiz_reduction"] > P.RUIZ_MIN_REDUCTION


def test_pivot_guard_skips_the_fp32_solve():
    """min |pivot| < 1e-9 max |A| after the P-fast refactor: the fp32 factors
    are not used for a solve, the chain goes straight to fp64 factors."""
    _Stub.built = []
    c = L.mna("c0", seed=0)
    x, rep = P.solve_adaptive(c, batch_factory=_Stub({"fast": "tiny-pivot", "safe": 1e-11, "hard": 1e-11}))
    a = rep["attempts"]
    assert rep["path"] == "P-safe" and len(a) == 2
    assert a[0]["pivot_guard"] and a[0]["min_pivot"] == 1e-30 and "rel_res" not in a[0]
    _Stub.built = []
    x, rep = P.solve_adaptive(c, batch_factory=_Stub({"fast": 1e-11, "safe": 1e-11, "hard": 1e-11}))
    assert rep["path"] == "P-fast" and "pivot_guard" not in rep["attempts"][0]
