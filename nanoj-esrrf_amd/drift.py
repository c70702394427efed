"""Drift of SR windows for the vibration correction (XYShiftCalculator.java:12-47), on
``esrrf_drift`` (include/esrrf.h; see there for the cross-correlation and peak fit used in
place of NanoJ-Core's, which the reference does not contain)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C


class XYShiftCalculator:
    """``new XYShiftCalculator(imp, prof)`` (:21-23) over a [n][h][w] stack;
    ``calculateShiftArray(indexStart, nFrameForSRRF)`` (:26-47) sets ``shiftX`` and
    ``shiftY`` of frames indexStart .. indexStart + nFrameForSRRF - 1 (ImageJ's 1-based
    slice numbers, :27,34) against frame indexStart, in the sign loadDriftXYGPUbuffer
    takes (a frame whose content moved by +d gets +d)."""

    def __init__(self, stack, prof=None, device=-1):
        self.ims = np.asarray(stack)
        if self.ims.ndim != 3:
            raise ValueError("stack must be [n][h][w]")
        n, h, w = self.ims.shape
        self.L = C.lib()
        hd = ctypes.c_void_p()
        C.check(self.L.esrrf_drift_create(int(w), int(h), int(device), ctypes.byref(hd)))
        self.h = hd
        self.shiftX = self.shiftY = None

    def calculateShiftArray(self, indexStart, nFrameForSRRF):
        s0, n = int(indexStart) - 1, int(nFrameForSRRF)
        if s0 < 0 or n < 1 or s0 + n > self.ims.shape[0]:
            raise ValueError("frames indexStart .. indexStart + nFrameForSRRF - 1 must exist")
        sub = np.ascontiguousarray(self.ims[s0:s0 + n], dtype=np.float32)
        sx = np.empty(n, np.float32)
        sy = np.empty(n, np.float32)
        rc = self.L.esrrf_drift_run(self.h, C.f32ptr(sub), n, 0, C.f32ptr(sx), C.f32ptr(sy))
        if rc != C.ESRRF_OK:
            m = self.L.esrrf_drift_last_error(self.h)
            raise C.EsrrfError(rc, m.decode() if m else "")
        self.shiftX, self.shiftY = sx, sy
        return sx, sy

    def close(self):
        if getattr(self, "h", None):
            self.L.esrrf_drift_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
