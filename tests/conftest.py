import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def oracle():
    from tests import oracle_lib
    return oracle_lib.load()


@pytest.fixture(scope="session")
def engine():
    from egsnrc_amd import _lib
    L = _lib.lib()
    import ctypes
    n = ctypes.c_int32(0)
    L.egsgpu_device_count(ctypes.byref(n))
    if n.value < 1:
        pytest.fail("GPU test run without a visible HIP device")
    return _lib
