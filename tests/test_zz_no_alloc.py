"""esrrf_frc_run_device_noalloc: the FRC of two device images with every buffer in the
library's static device scratch array, against the handle-based esrrf_frc_run_device on the
same images -- two calls in a row, the second smaller, so the second reuses the scratch
that the first left behind.  The ring sums accumulate in double with atomic adds, so two
runs of either entry point agree to rounding, not bit for bit: the float32 curve within
1e-6 and the resolution within 1e-9 relative."""
import ctypes
import math

import numpy as np
import pytest

from conftest import hip_runtime

pytestmark = pytest.mark.gpu


def sr_pair(h, w, seed):
    """Two noisy views of one band-limited structure (an odd/even-frame pair)."""
    rng = np.random.default_rng(seed)
    ky = np.fft.fftfreq(h)[:, None]
    kx = np.fft.fftfreq(w)[None, :]
    base = np.real(np.fft.ifft2(np.fft.fft2(rng.normal(size=(h, w))) *
                                np.exp(-(kx ** 2 + ky ** 2) / (2 * 0.08 ** 2))))
    base = 10 * base / np.abs(base).max()
    a = base + rng.normal(0, 0.3, (h, w))
    b = base + rng.normal(0, 0.3, (h, w))
    return a.astype(np.float32), b.astype(np.float32)


def both_entry_points(pkg, hip, a, b):
    """(resolution, curve) of esrrf_frc_run_device and of esrrf_frc_run_device_noalloc on
    the same two device images."""
    L = pkg._capi.lib()
    h, w = a.shape
    d = ctypes.c_void_p()
    nbytes = a.nbytes
    assert hip.hipMalloc(ctypes.byref(d), 2 * nbytes) == 0
    db = ctypes.c_void_p(d.value + nbytes)
    fp = ctypes.POINTER(ctypes.c_float)
    try:
        assert hip.hipMemcpy(d, a.ctypes.data, nbytes, 1) == 0      # host to device
        assert hip.hipMemcpy(db, b.ctypes.data, nbytes, 1) == 0
        q = ctypes.c_void_p()
        assert L.esrrf_frc_create(w, h, 0, ctypes.byref(q)) == 0
        nr, n = ctypes.c_int(), ctypes.c_int()
        L.esrrf_frc_rings(q, ctypes.byref(nr), ctypes.byref(n))
        exp_res = ctypes.c_double()
        exp_curve = np.empty(nr.value, np.float32)
        rc = L.esrrf_frc_run_device(q, d, db, ctypes.byref(exp_res),
                                    exp_curve.ctypes.data_as(fp))
        L.esrrf_frc_destroy(q)
        assert rc == 0, rc
        got_res, got_nr = ctypes.c_double(), ctypes.c_int()
        got_curve = np.full(4096, np.nan, np.float32)
        rc = L.esrrf_frc_run_device_noalloc(0, None, d, db, w, h, ctypes.byref(got_res),
                                            got_curve.ctypes.data_as(fp),
                                            ctypes.byref(got_nr))
        assert rc == 0, rc
        assert got_nr.value == nr.value
        return (exp_res.value, exp_curve), (got_res.value, got_curve[:got_nr.value].copy())
    finally:
        hip.hipFree(d)


def test_noalloc_frc_matches_handle_twice_second_smaller(pkg):
    hip = hip_runtime()
    for (h, w), seed in (((1280, 1280), 1), ((300, 220), 2)):
        a, b = sr_pair(h, w, seed)
        (exp_res, exp_curve), (got_res, got_curve) = both_entry_points(pkg, hip, a, b)
        np.testing.assert_allclose(got_curve, exp_curve, rtol=0, atol=1e-6)
        assert math.isfinite(exp_res)
        assert math.isclose(got_res, exp_res, rel_tol=1e-9), (h, w, got_res, exp_res)


def test_noalloc_frc_refuses_transforms_beyond_its_scratch(pkg):
    L = pkg._capi.lib()
    res = ctypes.c_double()
    nr = ctypes.c_int()
    dummy = ctypes.c_void_p(16)   # never read: the size check comes first
    ESRRF_ERR_UNSUPPORTED, ESRRF_ERR_INVALID = 5, 1
    assert L.esrrf_frc_run_device_noalloc(0, None, dummy, dummy, 2049, 2049, ctypes.byref(res),
                                          None, ctypes.byref(nr)) == ESRRF_ERR_UNSUPPORTED
    assert L.esrrf_frc_run_device_noalloc(0, None, None, dummy, 64, 64, ctypes.byref(res),
                                          None, ctypes.byref(nr)) == ESRRF_ERR_INVALID
