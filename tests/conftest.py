"""Test configuration: paths to the package and the oracle, the `gpu` marker, shared helpers."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "heatflow-mpm_amd"), os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built libhfm.so")


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    return float(num / den) if den > 0 else float(num)


@pytest.fixture(scope="session")
def hfm_lib():
    """libhfm.so loaded through ctypes (fails loudly if it is not built)."""
    from heatflow_mpm import _abi
    return _abi.load()


@pytest.fixture(scope="session")
def gpu(hfm_lib):
    n = hfm_lib.hfm_device_count()
    if n < 1:
        pytest.fail("no HIP device visible to libhfm.so")
    return n
