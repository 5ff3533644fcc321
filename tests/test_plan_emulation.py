"""The static plan as the device kernels use it, emulated on the CPU.

`plan_emulation.cpp` (this directory) is compiled with the host sources of the
library and run on generated circuits.
* The refactor's segment updates are emulated lane by lane, with every
  64-dependency chunk applied in parts cut at random ready prefixes, as the
  streaming waits cut them.  Every index is checked, and the result must match
  a per-dependency update.
* The forward and backward trisolve task lists, with their row and triangle
  packets and the 64-row block tasks of wide supernodes, are run in ticket
  order.  Every wait must already be satisfied, and the solution must match
  plain substitution.
* Three planted mistakes must be caught: a reversed task order, a lost map
  advance and two blocks of a wide supernode swapped.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HOST = os.path.join(os.path.dirname(HERE), "luspice_amd", "csrc", "host")


@pytest.fixture(scope="module")
def emulator(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("plan") / "plan_emulation")
    srcs = [os.path.join(HERE, "plan_emulation.cpp")] + \
        [os.path.join(HOST, f) for f in ("symbolic.cpp", "mna.cpp", "csc.cpp", "colamd.cpp")]
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", HOST, "-o", exe] + srcs, check=True)
    return exe


def _run(exe, *args):
    return subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=300)


@pytest.mark.parametrize("kind,tail,seed", [(0, 0, 1), (0, 300, 2), (3000, 0, 3), (3000, 1000, 4),
                                            (20000, 4000, 5), (20000, 0, 6)])
def test_plan_emulation(emulator, kind, tail, seed):
    r = _run(emulator, kind, tail, seed)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr


@pytest.mark.parametrize("plant,what,kind,tail", [(1, "fwd wait", 3000, 1000), (2, "map cursor", 3000, 1000),
                                                  (3, "fwd block wait", 3000, 0)])
def test_plan_emulation_catches_planted_mistakes(emulator, plant, what, kind, tail):
    r = _run(emulator, kind, tail, 4, plant)
    assert r.returncode != 0 and what in r.stdout, r.stdout
