"""Host phase of the batch API on the CPU (no device calls): lsp_batch_adopt
moves prepared circuits into a new batch without a second symbolic phase,
in the order picked, drops the circuits not picked and consumes the sources;
misuse is rejected."""
import numpy as np
import pytest

import luspice_amd as L


def _infos(B):
    return [B.info(i) for i in range(len(B.circuits))]


def test_adopt_moves_prepared_circuits():
    cs = [L.mna("sized", n=n, seed=30 + i) for i, n in enumerate([300, 520, 777])]
    A = L.Batch(cs[:2], upload=False)
    Bb = L.Batch(cs[2], upload=False)
    want = {c.n: info for c, info in zip(cs, _infos(A) + _infos(Bb))}
    est_parts = A.estimate_bytes() + Bb.estimate_bytes()
    M = L.Batch.adopt([(Bb, 0), (A, 0)], upload=False)
    assert A._h is None and Bb._h is None                 # consumed
    assert [c.n for c in M.circuits] == [777, 300] and M.total_n == 1077
    for c, info in zip(M.circuits, _infos(M)):
        assert info == want[c.n]                          # same pattern, levels, plan sizes
    assert np.array_equal(M.voff, [0, 777, 1077])
    assert M.estimate_bytes() <= est_parts                # circuit 1 of A was dropped
    M.close()


def test_adopt_rejects_a_circuit_picked_twice_and_bad_indices():
    c = L.mna("sized", n=200, seed=3)
    A = L.Batch(c, upload=False)
    with pytest.raises(RuntimeError):
        L.Batch.adopt([(A, 0), (A, 0)], upload=False)
    assert A._h is None                                   # consumed even on failure
    B = L.Batch(c, upload=False)
    with pytest.raises(RuntimeError):
        L.Batch.adopt([(B, 1)], upload=False)
