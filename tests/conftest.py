"""Test configuration: the `gpu` marker and import paths.

`-m "not gpu"` runs the oracle pins, the host-side operator algebra and the
library's exported symbols on CPU; `-m gpu` runs the device parity tests.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "neuralquantum.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libnqhip.so")


@pytest.fixture(scope="session")
def ref():
    from oracle import nqref
    return nqref.Ref()


@pytest.fixture(scope="session")
def gpu_ok():
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        if hip.hipGetDeviceCount(ctypes.byref(n)) != 0 or n.value == 0:
            pytest.fail("no HIP device visible")
    except OSError as e:
        pytest.fail(f"HIP runtime missing: {e}")
    return True
