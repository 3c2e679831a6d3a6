import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def rel_l2(a, b):
    """||a - b||_2 / ||b||_2 over all entries (0 when both are zero)."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    na = np.linalg.norm(a - b)
    if nb == 0.0:
        return na
    return na / nb


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle_py

    oracle_py.oracle_library()
    return oracle_py


@pytest.fixture(scope="session")
def gpu():
    """The HIP library, loaded; fails (not skips) when no GPU is visible."""
    import ctypes

    from sphnj_amd import sim

    lib = sim.library()
    n = ctypes.c_int(0)
    hip = ctypes.CDLL("libamdhip64.so")
    rc = hip.hipGetDeviceCount(ctypes.byref(n))
    if rc != 0 or n.value < 1:
        pytest.fail("GPU test selected but no GPU is visible")
    return lib
