"""Shared fixtures for the vlp4d tests.

Markers: ``gpu`` for tests that need an MI355X (they call libvlp4d.so through
its C-ABI); everything else runs on the CPU and checks the oracle, the host
logic and the library's exported surface.
"""
import importlib.util
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
DECKS = os.path.join(GOLDEN, "decks")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X GPU (HIP library calls)")


def load_pkg():
    """Import the p3-miniapps_amd package (directory name is not an identifier)."""
    name = "p3_miniapps_amd"
    if name in sys.modules:
        return sys.modules[name]
    path = os.path.join(ROOT, "p3-miniapps_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def vlp():
    return load_pkg()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.load()
    return oracle


def sld_box(n, vmax=6.0, dt=0.05, nbiter=10, idcase=10, lx=4 * math.pi):
    """(n, minPhy, maxPhy, idcase, dt, nbiter) of an SLD10-style deck."""
    n = list(n) if hasattr(n, "__len__") else [n] * 4
    return n, [0.0, 0.0, -vmax, -vmax], [lx, lx, vmax, vmax], idcase, dt, nbiter


def relmax(a, b):
    """max|a-b| / max|b| (the parity metric: relative max-norm)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))
