import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "lespmv_golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: larger GPU cases")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.fixture(scope="session")
def gpu():
    """The HIP library, initialised on device 0. Fails loudly if it cannot load."""
    from sspmv_amd import lib
    lib.load()
    lib.check(lib.load().lespmv_init(0))
    return lib


def rel2(a, b):
    """Relative 2-norm error ||a-b|| / ||b|| (b the oracle)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


# tolerances stated by the task: relative 2-norm error vs the FP64 oracle
TOL = {np.dtype(np.float64): 1e-12, np.dtype(np.float32): 1e-5}
