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
        key = ("safe" if opts.get("factor_shift") else "hard") if opts.get("pivot_scaled") else "fast"
        _Stub.built.append((key, opts))
        s = self.script[key]
        if s == "singular":
            raise RuntimeError("symbolic factor: singular column 3")
        return _StubBatch(circuit.n, s)


class _StubBatch:
    def __init__(self, n, s):
        self.n, self.s = n, s

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
