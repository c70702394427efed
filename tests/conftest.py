import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

import __graft_entry__  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def pkg():
    return __graft_entry__.load()


@pytest.fixture(scope="session")
def O():
    from oracle import oracle
    oracle.lib()
    return oracle


def normalised_close(got, exp, rel=1e-4, abs_=1e-5):
    """Tolerance of the port: |a-b| <= abs + rel*|b| after dividing both images by the
    oracle image's max |value| (per channel).  Returns (ok, worst excess, index)."""
    got = np.asarray(got, np.float64)
    exp = np.asarray(exp, np.float64)
    scale = max(float(np.abs(exp).max()), 1e-30)
    a, b = got / scale, exp / scale
    excess = np.abs(a - b) - (abs_ + rel * np.abs(b))
    i = int(np.argmax(excess))
    return bool(excess.flat[i] <= 0), float(excess.flat[i]), np.unravel_index(i, excess.shape)


def hip_runtime():
    """The HIP runtime libesrrf.so runs on, through ctypes: device and pinned host buffers
    for the tests of the *_device entry points (torch ships a second HIP runtime of its
    own, which tests keep out of the process)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t,
                                  ctypes.c_uint]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipHostFree.argtypes = [ctypes.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    return hip
