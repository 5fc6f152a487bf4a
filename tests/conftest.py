"""Shared fixtures.  Tests marked ``gpu`` need an MI355X and the built
libmpmx.so; the rest run on the CPU (oracle pinning, host logic, ABI)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libmpmx.so")


@pytest.fixture(scope="session")
def o64():
    from oracle import Oracle
    return Oracle(64)


@pytest.fixture(scope="session")
def o32():
    from oracle import Oracle
    return Oracle(32)


def rel_err(a, b):
    """max |a - b| / max |b| over the whole array (error relative to the array's scale)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    if scale == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / scale)
