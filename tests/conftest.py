"""pytest configuration: the `gpu` marker and shared fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def c0():
    import luspice_amd as L
    return L.mna("c0", seed=0)


@pytest.fixture(scope="session")
def c0_q(c0):
    import luspice_amd as L
    return L.colamd(c0)
