import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture
def make_solver():
    """Factory for GPU solver contexts (closed after the test)."""
    from stokes_to_amd.steady import FluidSteadyState
    made = []

    def _make(dim, n, **kw):
        s = FluidSteadyState(dim, n, **kw)
        made.append(s)
        return s

    yield _make
    for s in made:
        s.close()


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.linalg.norm(a - b)
    n = max(np.linalg.norm(b), 1e-300)
    return d / n
