import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import build as ob
    from oracle import oracle as O
    if not os.path.exists(O.LIB):
        ob.build()
    return O


def _hip_device_present():
    try:
        import ctypes
        h = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return h.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def gpu():
    """The product library on a real device; fails (not skips) without one."""
    from phi_eng_amd import ffi
    ffi.lib()
    if not _hip_device_present():
        pytest.fail("no HIP device visible to a test marked gpu")
    return ffi
