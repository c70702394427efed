"""SQUIRREL RSE of the eSRRF AVG image and its gradient (ErrorMapLiveSRRF.java on the
sweep's images, gui/ParametersSweep_.java:473-476), on the C ABI.

``RSE(width, height, M, n_frames)`` loads a stack once; ``value_and_grad(fwhm, S, sigma)``
returns the RSE and dRSE/d(fwhm, S, sigma); ``best_sigma(fwhm, S)`` minimises over sigma
as ErrorMapLiveSRRF.optimise does (Brent on [0, 4/2.35482 M]); ``descend`` runs projected
Adam on the three parameters.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C


class RSE:
    def __init__(self, width, height, magnification, n_frames, crop_size=3, int_weighting=True,
                 mp_correction=True, sigma_max=0.0, device=-1):
        self.L = C.lib()
        p = C.RseParams()
        p.width, p.height, p.magnification = int(width), int(height), int(magnification)
        p.n_frames, p.crop_size = int(n_frames), int(crop_size)
        p.int_weighting = 1 if int_weighting else 0
        p.do_mp_map_correction = 1 if mp_correction else 0
        p.sigma_max = float(sigma_max)
        p.device = int(device)
        h = ctypes.c_void_p()
        C.check(self.L.esrrf_rse_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        self.sigma_max = float(sigma_max) if sigma_max > 0 else 4.0 / 2.35482 * magnification

    def _check(self, rc):
        if rc != C.ESRRF_OK:
            m = self.L.esrrf_rse_last_error(self.h)
            raise C.EsrrfError(rc, m.decode() if m else "")

    def load(self, stack):
        a = np.asarray(stack)
        dt = C.ESRRF_U16 if a.dtype == np.uint16 else C.ESRRF_F32
        if dt == C.ESRRF_F32:
            a = a.astype(np.float32, copy=False)
        a = np.ascontiguousarray(a)
        self._check(self.L.esrrf_rse_load(self.h, a.ctypes.data_as(ctypes.c_void_p), dt,
                                          a.shape[0]))

    def value_and_grad(self, fwhm, sensitivity, sigma, grad=True):
        r = ctypes.c_double()
        g = (ctypes.c_double * 3)()
        self._check(self.L.esrrf_rse_value_and_grad(self.h, fwhm, sensitivity, sigma,
                                                    ctypes.byref(r), g if grad else None))
        return (r.value, np.array(list(g))) if grad else r.value

    def value(self, fwhm, sensitivity, sigma):
        return self.value_and_grad(fwhm, sensitivity, sigma, grad=False)

    def best_sigma(self, fwhm, sensitivity):
        s, r, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self._check(self.L.esrrf_rse_best_sigma(self.h, fwhm, sensitivity, ctypes.byref(s),
                                                ctypes.byref(r), ctypes.byref(n)))
        return s.value, r.value, n.value

    def descend(self, x0, steps=30, lr=None, lo=None, hi=None):
        f3 = lambda v: None if v is None else (ctypes.c_float * 3)(*[float(t) for t in v])
        traj = np.zeros((steps + 1, 4), np.float64)
        self._check(self.L.esrrf_rse_descend(
            self.h, f3(x0), int(steps), f3(lr), f3(lo), f3(hi),
            traj.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return traj

    def counters(self):
        e, p = ctypes.c_longlong(), ctypes.c_longlong()
        self._check(self.L.esrrf_rse_counters(self.h, ctypes.byref(e), ctypes.byref(p)))
        return e.value, p.value

    def close(self):
        if self.h:
            self.L.esrrf_rse_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
