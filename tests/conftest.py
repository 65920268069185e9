"""Shared fixtures.  Tests marked ``gpu`` need an MI355X and the built
libglsm_fluid.so; everything else runs on the CPU (the oracle is built on
demand with g++)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.build()
    O.set_threads(min(16, os.cpu_count() or 1))
    return O


@pytest.fixture(scope="session")
def glsm():
    """The HIP library, initialised on device 0.  Fails (never skips) when
    the library or the GPU is missing: GPU tests must run the native path."""
    import glsm_fluid_3d_amd as G

    G.load()
    G.init(0)
    return G


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a - b))
