"""SQUIRREL error map of an SR image (ErrorMapLiveSRRF.java), the FRC resolution and the
QnR map (gui/GetQnRmap_.java) over the C ABI.

``ErrorMapLiveSRRF`` keeps the reference class's constructor, ``optimise(ref, sr)`` and
``calculateErrorMap()``, and its public results (``globalRMSE``, ``globalPPMCC``, ``fpSRC``,
``fpErrorMap``); the sigma search, blur, fit and maps run on the GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C


class ErrorMapLiveSRRF:
    """Mirror of ErrorMapLiveSRRF (ErrorMapLiveSRRF.java:19-78).  ``imp_shape`` is the
    (height, width) of the raw image the reference's constructor receives."""

    def __init__(self, imp_shape, magnification, fixSigma=False, fixedSigma=0.0,
                 cropBorder=True, cropSize=3, device=-1):
        self.L = C.lib()
        h, w = int(imp_shape[0]), int(imp_shape[1])
        p = C.SquirrelParams()
        p.width, p.height, p.magnification = w, h, int(magnification)
        p.crop_border = 1 if cropBorder else 0
        p.crop_size = int(cropSize)
        p.fix_sigma = 1 if fixSigma else 0
        p.fixed_sigma = float(fixedSigma)
        p.device = int(device)
        hd = ctypes.c_void_p()
        C.check(self.L.esrrf_squirrel_create(ctypes.byref(p), ctypes.byref(hd)))
        self.h = hd
        self.magnification = int(magnification)
        self.full = (h * int(magnification), w * int(magnification))
        wc, hc = ctypes.c_int(), ctypes.c_int()
        self._check(self.L.esrrf_squirrel_dims(self.h, ctypes.byref(wc), ctypes.byref(hc)))
        self.width, self.height = wc.value, hc.value
        self._res = None

    def _check(self, rc):
        if rc != C.ESRRF_OK:
            m = self.L.esrrf_squirrel_last_error(self.h)
            raise C.EsrrfError(rc, m.decode() if m else "")

    def optimise(self, impRef, impSR):
        """:81-153: sigma (Brent or fixed) and alpha, beta for this (ref, SR) pair."""
        ref = np.ascontiguousarray(impRef, dtype=np.float32)
        sr = np.ascontiguousarray(impSR, dtype=np.float32)
        if ref.shape != self.full or sr.shape != self.full:
            raise ValueError(f"images must be {self.full}, got {ref.shape} and {sr.shape}")
        emap = np.empty((self.height, self.width), np.float32)
        rsc = np.empty_like(emap)
        r = C.SquirrelResult()
        self._check(self.L.esrrf_squirrel_run(self.h, C.f32ptr(ref), C.f32ptr(sr),
                                              ctypes.byref(r), C.f32ptr(emap), C.f32ptr(rsc)))
        self._res = r
        self._maps = (emap, rsc)
        self.sigma_linear, self.alpha, self.beta = r.sigma, r.alpha, r.beta
        self.overblur = bool(r.overblur)

    def calculateErrorMap(self):
        """:156-190: RSC, globalRMSE, globalPPMCC and the error map of the last optimise."""
        if self._res is None:
            raise RuntimeError("optimise() first")
        self.fpErrorMap, self.fpSRC = self._maps
        self.globalRMSE = self._res.rse
        self.globalPPMCC = self._res.rsp

    def close(self):
        if getattr(self, "h", None):
            self.L.esrrf_squirrel_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def remove_nans(img, radius=2.0):
    """One pass of ImageJ's "Remove NaNs..." (see esrrf_remove_nans) on a copy of a 2-D map."""
    a = np.array(img, dtype=np.float32, order="C", copy=True)
    if a.ndim != 2:
        raise ValueError("map must be 2-D")
    C.check(C.lib().esrrf_remove_nans(C.f32ptr(a), a.shape[1], a.shape[0], float(radius)))
    return a


def qnr_map(frc_map, rsp_map, auto_min_max=True, res_min=0.05, res_max=0.2, removeNaNs=False):
    """GetQnRmap_ (gui/GetQnRmap_.java:179-258): logistic conversion of the FRC resolution
    map, QnR = 2 L RSP / (L + RSP), and its first maximum.  Returns (qnr, best, (x, y))
    with x the fwhm index and y the sensitivity index of a [n_sensitivity][n_fwhm] map.
    removeNaNs runs "Remove NaNs..." (radius 2) as the plugin does (:69-77): three passes on
    the RSP map, two on the FRC map, then a fourth on the RSP map."""
    frc = np.ascontiguousarray(frc_map, dtype=np.float32)
    rsp = np.ascontiguousarray(rsp_map, dtype=np.float32)
    if frc.shape != rsp.shape:
        raise ValueError("FRC and RSP maps must have the same shape")
    if removeNaNs:
        for _ in range(3):
            rsp = remove_nans(rsp, 2.0)
        for _ in range(2):
            frc = remove_nans(frc, 2.0)
        rsp = remove_nans(rsp, 2.0)
    q = np.empty_like(frc)
    best = ctypes.c_float()
    idx = ctypes.c_long()
    C.check(C.lib().esrrf_qnr_map(C.f32ptr(frc), C.f32ptr(rsp), frc.size,
                                  1 if auto_min_max else 0, float(res_min), float(res_max),
                                  C.f32ptr(q), ctypes.byref(best), ctypes.byref(idx)))
    w = frc.shape[-1] if frc.ndim >= 1 else 1
    return q, best.value, (idx.value % w, idx.value // w)


class FRC:
    """Fourier ring correlation of two images of one size (the sweep's
    ``FRC.calculateFireNumber(ip1, ip2, FIXED_1_OVER_7)``; NanoJ-Core's class is not part
    of the reference, see esrrf_frc_create for the definition used)."""

    def __init__(self, width, height, device=-1):
        self.L = C.lib()
        h = ctypes.c_void_p()
        C.check(self.L.esrrf_frc_create(int(width), int(height), int(device), ctypes.byref(h)))
        self.h = h
        self.shape = (int(height), int(width))
        nr, n = ctypes.c_int(), ctypes.c_int()
        self.L.esrrf_frc_rings(self.h, ctypes.byref(nr), ctypes.byref(n))
        self.n_rings, self.transform_size = nr.value, n.value

    def calculateFireNumber(self, ip1, ip2):
        """Resolution in pixels; the curve of the last call is in ``self.curve``."""
        a = np.ascontiguousarray(ip1, dtype=np.float32)
        b = np.ascontiguousarray(ip2, dtype=np.float32)
        if a.shape != self.shape or b.shape != self.shape:
            raise ValueError(f"images must be {self.shape}")
        res = ctypes.c_double()
        self.curve = np.empty(self.n_rings, np.float32)
        rc = self.L.esrrf_frc_run(self.h, C.f32ptr(a), C.f32ptr(b), ctypes.byref(res),
                                  C.f32ptr(self.curve))
        if rc != C.ESRRF_OK:
            m = self.L.esrrf_frc_last_error(self.h)
            raise C.EsrrfError(rc, m.decode() if m else "")
        return res.value

    def close(self):
        if getattr(self, "h", None):
            self.L.esrrf_frc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
