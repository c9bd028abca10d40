import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

# Tests that hand device buffers to the library allocate them with torch.
# torch must load its HIP runtime before libbbrcit_hip.so pulls in the system
# one; in the other order torch finds no device in the same process.
try:
    import torch  # noqa: F401
except Exception:  # torch is optional for the CPU tests
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle.oracle import REF_LIB, Reference

    if not REF_LIB.exists():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return Reference()


@pytest.fixture(scope="session")
def bk():
    import bbrcit_kde_amd

    bbrcit_kde_amd.load_library()
    return bbrcit_kde_amd


def rel_errors(got, want, floor):
    """Relative errors where |want| >= floor, absolute/floor elsewhere."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    denom = np.maximum(np.abs(want), floor)
    return np.abs(got - want) / denom
