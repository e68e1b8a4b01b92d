import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.build()
    return pyoracle


def gpu_available() -> bool:
    """Counts HIP devices through the runtime library (no torch import)."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        try:
            hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        except OSError:
            return False
    n = ctypes.c_int(0)
    return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0


@pytest.fixture(scope="session")
def gpu():
    """The engine library on a real GPU; a GPU test without one is an error,
    not a skip, so a silent CPU run can never pass for a GPU run."""
    if not gpu_available():
        pytest.fail("no GPU visible: -m gpu tests must run on an MI355X")
    from jmmonedmc_amd import _native
    return _native.lib()
