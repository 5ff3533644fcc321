"""The persistent kernels' split task queue, simulated on the CPU.

`next_task` (luspice_amd/csrc/kernels/common.hpp) spreads a task list over nq
counters: task t belongs to counter t % nq, wave w of every 4-wave workgroup
takes its tickets from counter w % nq, and ticket k of counter q is task
k * nq + q.  Tasks wait only on lower tasks.  The claim is that no schedule
deadlocks, whatever the occupancy: every counter has a wave in every
workgroup, so even one resident workgroup drains the queue.

The simulation runs waves in random order on random dependency graphs (long
chains, as the refactor's column chains, plus random earlier tasks), with
workgroups made resident only as earlier ones finish, and requires every task
to complete.  A planted variant that gives each workgroup a single counter must
be caught deadlocking with one resident workgroup.
"""
import random

import pytest

WAVES = 4   # waves per workgroup, as the refactor and trisolve grids launch


def _graph(rng, n):
    deps = [[] for _ in range(n)]
    for t in range(1, n):
        if rng.random() < 0.6:
            deps[t].append(t - 1)                      # chain link
        for _ in range(rng.randrange(3)):
            deps[t].append(rng.randrange(t))           # any earlier task
    return deps


def _simulate(deps, nq, groups, resident, per_wave_counter=True, seed=0):
    """Returns True when every task finished, False on a deadlock."""
    rng = random.Random(seed)
    n = len(deps)
    ctr = [0] * nq
    done = [False] * n
    queue = list(range(groups))                        # workgroups not yet resident
    live = {}                                          # (group, wave) -> held task, None = idle
    exited = {}

    def admit():
        while queue and len({g for g, _ in live}) < resident:
            g = queue.pop(0)
            exited[g] = 0
            for w in range(WAVES):
                live[(g, w)] = None

    admit()
    while live:
        progressed = False
        for key in rng.sample(list(live), len(live)):
            g, w = key
            task = live[key]
            if task is None:
                q = (w if per_wave_counter else g) % nq
                t = ctr[q] * nq + q
                ctr[q] += 1
                if t >= n:                             # past the end: the wave leaves
                    del live[key]
                    exited[g] += 1
                    if exited[g] == WAVES:
                        admit()
                else:
                    live[key] = t
                progressed = True
            elif all(done[d] for d in deps[task]):
                done[task] = True
                live[key] = None
                progressed = True
        if not progressed:
            return False
    return all(done)


@pytest.mark.parametrize("nq", [1, 2, 4])
@pytest.mark.parametrize("resident", [1, 3, 16])
def test_split_queue_never_deadlocks(nq, resident):
    rng = random.Random(100 * nq + resident)
    for trial in range(4):
        deps = _graph(rng, 400)
        assert _simulate(deps, nq, groups=16, resident=resident, seed=trial), (nq, resident, trial)


def test_planted_counter_per_workgroup_deadlocks():
    # one counter per workgroup: with one resident workgroup, counters 1..3
    # are never served and the chain through them stops
    deps = [[t - 1] if t else [] for t in range(64)]
    assert not _simulate(deps, 4, groups=8, resident=1, per_wave_counter=False)
    assert _simulate(deps, 4, groups=8, resident=1, per_wave_counter=True)
