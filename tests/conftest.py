import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def pg():
    """The pygranite-compatible module (native build required)."""
    import pygranite_amd
    return pygranite_amd


@pytest.fixture(scope="session")
def gpu_available():
    from pygranite_amd import capi
    n = capi.device_count()
    if n == 0:
        pytest.fail("GPU test selected but no HIP device is visible")
    return n
