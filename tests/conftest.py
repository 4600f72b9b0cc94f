import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


_has_gpu = None


def has_gpu() -> bool:
    global _has_gpu
    if _has_gpu is None:
        import ctypes
        try:
            hip = ctypes.CDLL("libamdhip64.so")
            n = ctypes.c_int(0)
            _has_gpu = hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
        except OSError:
            _has_gpu = False
    return _has_gpu


@pytest.fixture
def gpu():
    if not has_gpu():
        pytest.skip("no HIP device in this environment")
    return True
