import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def golden_kernels():
    return np.load(os.path.join(GOLDEN, "kernels.npz"))


@pytest.fixture(scope="session")
def golden_varifold():
    return np.load(os.path.join(GOLDEN, "varifold.npz"))


def rel_err(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a)))


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on a machine with a GPU; fails (never skips) when the
    library is missing, so a GPU run can't pass on a fallback."""
    import lddmm_amd
    L = lddmm_amd.lib()
    import ctypes
    n = ctypes.c_int(0)
    L.lddmm_device_count(ctypes.byref(n))
    assert n.value >= 1, "no HIP device visible"
    return lddmm_amd
