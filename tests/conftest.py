import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an AMD GPU (runs the HIP engine)')
    config.addinivalue_line('markers', 'slow: longer CPU test')


def _has_gpu():
    """True when the HIP runtime reports a device (no torch import needed)."""
    import ctypes
    try:
        hip = ctypes.CDLL('libamdhip64.so')
    except OSError:
        try:
            hip = ctypes.CDLL('/opt/rocm/lib/libamdhip64.so')
        except OSError:
            return False
    n = ctypes.c_int(0)
    return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0


def pytest_collection_modifyitems(config, items):
    if _has_gpu():
        return
    skip = pytest.mark.skip(reason='no GPU visible')
    for item in items:
        if 'gpu' in item.keywords:
            item.add_marker(skip)
