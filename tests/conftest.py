import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through librblas_hip.so)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU gate: where the gpu tests run, the library must load and
    see a device."""
    from rblas_amd import device
    n = device.device_count()
    assert n >= 1, "gpu tests need a visible HIP device"
    return n
