"""Parameter sweep (gui/ParametersSweep_.java) on the fused device sequence of the C ABI.

``ParametersSweep`` takes the sweep's parameter axes as the GUI builds them (:720-747) and
returns, for each (magnification, frame count, sensitivity, fwhm), the four channels that
the reference's per-configuration ``LiveSRRF_CL.initialise`` / ``calculateSRRF`` /
``readSRRFbuffer`` produce (:452-458, :800-844), over the first nf frames of the stack.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi as C


def _java_idiv(a: int, b: int) -> int:
    """Java int division (truncation toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def sweep_axes(fwhm0, dfwhm, n_fwhm, S0, dS, n_S, nf0, dnf, n_nf, n_slices, frc=False):
    """The parameter arrays of ParametersSweep_.grabSettings (:720-747), Java arithmetic:
    fwhm in float, sensitivity and frame counts in int, frame counts capped by the stack."""
    fwhm = [float(np.float32(np.float32(fwhm0) + np.float32(i) * np.float32(dfwhm)))
            for i in range(n_fwhm)]
    sens = [int(S0) + i * int(dS) for i in range(n_S)]
    avail = (n_slices // 2 if frc else n_slices) - int(nf0)
    n_use = min(_java_idiv(avail, int(dnf)) + 1, int(n_nf))
    nfs = [int(nf0) + i * int(dnf) for i in range(max(n_use, 0))]
    return fwhm, sens, nfs


class ParametersSweep:
    def __init__(self, width, height, magnifications, fwhm, sensitivity, n_frames,
                 int_weighting=True, mp_correction=True, autotune=False, use_graph=True,
                 device=-1, fht=False):
        """fht: the reference sweep's doFHTinterpolation (gui/ParametersSweep_.java:117,452):
        Fourier-upsampled frames, no MP correction."""
        self.L = C.lib()
        self.width, self.height = int(width), int(height)
        self.mags = np.ascontiguousarray(magnifications, dtype=np.int32)
        self.fwhm = np.ascontiguousarray(fwhm, dtype=np.float32)
        self.sens = np.ascontiguousarray(sensitivity, dtype=np.float32)
        self.nfs = np.ascontiguousarray(n_frames, dtype=np.int32)
        p = C.SweepParams()
        p.width, p.height = self.width, self.height
        ip = ctypes.POINTER(ctypes.c_int)
        fp = ctypes.POINTER(ctypes.c_float)
        p.magnifications, p.n_magnifications = self.mags.ctypes.data_as(ip), self.mags.size
        p.fwhm, p.n_fwhm = self.fwhm.ctypes.data_as(fp), self.fwhm.size
        p.sensitivity, p.n_sensitivity = self.sens.ctypes.data_as(fp), self.sens.size
        p.n_frames, p.n_n_frames = self.nfs.ctypes.data_as(ip), self.nfs.size
        p.int_weighting = 1 if int_weighting else 0
        p.do_mp_map_correction = 1 if mp_correction else 0
        p.autotune = 1 if autotune else 0
        p.use_graph = 1 if use_graph else 0
        p.device = int(device)
        p.do_fht_interpolation = 1 if fht else 0
        self.fht = bool(fht)
        h = ctypes.c_void_p()
        C.check(self.L.esrrf_sweep_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h

    def _check(self, rc):
        if rc != C.ESRRF_OK:
            m = self.L.esrrf_sweep_last_error(self.h)
            raise C.EsrrfError(rc, m.decode() if m else "")

    def run(self, stack):
        a = np.asarray(stack)
        dt = C.ESRRF_U16 if a.dtype == np.uint16 else C.ESRRF_F32
        if dt == C.ESRRF_F32:
            a = a.astype(np.float32, copy=False)
        a = np.ascontiguousarray(a)
        self._check(self.L.esrrf_sweep_run(self.h, a.ctypes.data_as(ctypes.c_void_p), dt,
                                           a.shape[0]))

    def run_device(self, ptr, dtype, n_frames):
        """Frames already in device memory (pointer), asynchronous on the sweep's stream."""
        self._check(self.L.esrrf_sweep_run_device(self.h, ctypes.c_void_p(ptr), dtype,
                                                  int(n_frames)))

    def set_stream(self, stream_ptr):
        self._check(self.L.esrrf_sweep_set_stream(self.h, ctypes.c_void_p(stream_ptr)))

    def read(self, i_mag, i_nf, i_sens, i_fwhm):
        M = int(self.mags[i_mag])
        out = np.empty((4, self.height * M, self.width * M), np.float32)
        self._check(self.L.esrrf_sweep_read(self.h, i_mag, i_nf, i_sens, i_fwhm, C.f32ptr(out)))
        return out

    def shape(self, i_mag, i_fwhm=0):
        """Launch shape of the RGC emit launch of one (magnification, fwhm)."""
        nw, ppw, tl, fpp = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = ctypes.c_float()
        self._check(self.L.esrrf_sweep_shape(self.h, i_mag, i_fwhm, ctypes.byref(nw),
                                             ctypes.byref(ppw), ctypes.byref(tl),
                                             ctypes.byref(fpp), ctypes.byref(ms)))
        return dict(nw=nw.value, ppw=ppw.value, taps_in_lds=tl.value,
                    frames_per_pass=fpp.value, ms=ms.value)

    def error_maps(self, channel="AVG", crop_size=3, fix_sigma=False, fixed_sigma=0.0):
        """The sweep's error mapping (gui/ParametersSweep_.java:465-530) for one channel
        ("AVG", "VAR" or "TAC2") against each configuration's INT image: RSE, RSP and RSF
        sigma arrays [n_magnifications][n_n_frames][n_sensitivity][n_fwhm]."""
        ch = {"AVG": 0, "VAR": 1, "TAC2": 2}[channel]
        shape = (self.mags.size, self.nfs.size, self.sens.size, self.fwhm.size)
        rse, rsp, sig = (np.empty(shape, np.float64) for _ in range(3))
        dp = ctypes.POINTER(ctypes.c_double)
        self._check(self.L.esrrf_sweep_squirrel(
            self.h, ch, int(crop_size), 1 if fix_sigma else 0, float(fixed_sigma),
            rse.ctypes.data_as(dp), rsp.ctypes.data_as(dp), sig.ctypes.data_as(dp)))
        return dict(rse=rse, rsp=rsp, sigma=sig)

    def frc_maps(self, even, channel="AVG"):
        """FRC resolution (SR pixels) of every configuration between this sweep, run on the
        odd frames, and ``even``, run on the even frames with the same parameters
        (gui/ParametersSweep_.java:403-447): [n_magnifications][n_n_frames][n_S][n_fwhm]."""
        ch = {"AVG": 0, "VAR": 1, "TAC2": 2}[channel]
        out = np.empty((self.mags.size, self.nfs.size, self.sens.size, self.fwhm.size),
                       np.float64)
        self._check(self.L.esrrf_sweep_frc(self.h, even.h, ch,
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def close(self):
        if self.h:
            self.L.esrrf_sweep_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def odd_even(stack, n_frames):
    """The odd and even frame sets of the sweep's FRC mode (gui/ParametersSweep_.java:808-816:
    frame f of a configuration is stack slice 2(f-1)+1, resp. 2f, 1-based)."""
    a = np.asarray(stack)
    n = int(max(n_frames))
    return a[0:2 * n:2], a[1:2 * n:2]
