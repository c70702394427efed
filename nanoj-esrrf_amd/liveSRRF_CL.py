"""Host-side mirror of the reference's engine class ``LiveSRRF_CL`` on the HIP C ABI.

Same public methods, argument meaning and read-back layout as
src/nanoj/liveSRRF/LiveSRRF_CL.java; each method cites the Java method it mirrors.
ImageJ stacks are numpy arrays here: a stack is [n][height][width] (uint16 or float32),
``imsSRRF`` is [4 * nPlanesM][heightM][widthM] (AVG, VAR, TAC2, INT slices).

Differences from the Java class, all deliberate:
* ``CLDevice`` becomes :class:`Device` (``getName()``), the only part callers use
  (gui/LiveSRRF_optimised_.java:157-188, gui/ParametersSweep_.java:127-202);
* the 3-D calibration table is passed as a mapping with the .njt column names
  (``"X-shift (pixels)"``, ``"Y-shift (pixels)"``, ``"Theta (degrees)"``, ``"ROI #"``,
  ``"Intensity scaling"``, ``"Background level (ADC)"``) instead of a file path, because the
  NanoJ table reader (nanoj.core.java.io.LoadNanoJTable) is not part of the reference;
* cancellation is a ``cancel`` callable polled between launch chunks instead of
  ``IJ.escapePressed()``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _capi as C


@dataclass
class Device:
    index: int
    name: str

    def getName(self) -> str:
        return self.name


class Engine:
    """Thin RAII owner of one ``esrrf_engine`` handle."""

    def __init__(self, params: C.Params):
        self.L = C.lib()
        h = ctypes.c_void_p()
        rc = self.L.esrrf_create(ctypes.byref(params), ctypes.byref(h))
        if rc != C.ESRRF_OK:
            raise C.EsrrfError(rc, "esrrf_create failed")
        self.h = h
        self.params = params
        wm, hm, npm, ws, hs, npl = (ctypes.c_int() for _ in range(6))
        C.check(self.L.esrrf_get_dims(self.h, *(ctypes.byref(v) for v in
                                                (wm, hm, npm, ws, hs, npl))), self.h)
        self.wM, self.hM, self.nPlanesM = wm.value, hm.value, npm.value
        self.widthS, self.heightS, self.nPlanes = ws.value, hs.value, npl.value

    def close(self):
        if getattr(self, "h", None):
            self.L.esrrf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        C.check(rc, self.h)

    def counters(self):
        out = (ctypes.c_int * 2)()
        self.check(self.L.esrrf_read_frame_counters(self.h, out))
        return [out[0], out[1]]

    def rgc_kernel(self) -> str:
        """Name of the RGC kernel the last calculate launched."""
        buf = ctypes.create_string_buffer(64)
        self.check(self.L.esrrf_rgc_kernel(self.h, buf, 64))
        return buf.value.decode()


class LiveSRRF_CL:
    """Mirror of LiveSRRF_CL (LiveSRRF_CL.java:29-1017)."""

    nReconstructions = 3          # :49
    gradientMag = 2               # :44

    def __init__(self, DEBUG: bool = False):   # :105-108
        self.DEBUG = DEBUG
        self.allCLdevices: list[Device] = []
        self.imsSRRF = None
        self.widthS = self.heightS = self.nPlanes = self.nPlanesM = 0
        self._eng: Engine | None = None
        self._calib = None

    # -- :111-155
    def checkDevices(self):
        L = C.lib()
        n = ctypes.c_int()
        C.check(L.esrrf_device_count(ctypes.byref(n)))
        if n.value == 0:
            raise RuntimeError("Something went wrong initializing HIP: no device.")
        self.allCLdevices = []
        for i in range(n.value):
            buf = ctypes.create_string_buffer(256)
            C.check(L.esrrf_device_name(i, buf, 256))
            self.allCLdevices.append(Device(i, buf.value.decode()))

    # -- :159-264
    def initialise(self, width, height, magnification, fwhm, sensitivity, nFramesOnGPU,
                   nFrameForSRRF, blockLength, chosenDevice, intWeighting, doMPmapCorrection,
                   doFHTinterpolation, calibTablePath3DSRRF=None, axialOffset=0.0,
                   pixelSize=1.0, axialMagnification=0):
        """The reference's initialise; ``axialMagnification`` (3-D only, no counterpart in
        the reference) sets the output planes per input plane, 0 meaning ``magnification``
        as live3DSRRF.cl has it."""
        self.release()
        p = C.Params()
        p.width, p.height, p.magnification = int(width), int(height), int(magnification)
        p.fwhm = float(fwhm)
        p.sensitivity = float(sensitivity)
        p.n_frames_on_gpu = int(nFramesOnGPU)
        p.n_frame_for_srrf = int(nFrameForSRRF)
        p.block_length = int(blockLength)
        p.device = -1 if chosenDevice is None else int(chosenDevice.index)
        p.int_weighting = 1 if intWeighting else 0
        p.do_mp_map_correction = 1 if doMPmapCorrection else 0
        p.do_fht_interpolation = 1 if doFHTinterpolation else 0
        self._calib = None
        if calibTablePath3DSRRF is not None:
            t = calibTablePath3DSRRF
            cols = {k: np.ascontiguousarray(np.asarray(t[k], dtype=np.float64)) for k in
                    ("X-shift (pixels)", "Y-shift (pixels)", "Theta (degrees)", "ROI #",
                     "Intensity scaling")}
            self._calib = cols
            self._bg = float(np.asarray(t.get("Background level (ADC)", [0.0]))[0])
            dp = ctypes.POINTER(ctypes.c_double)
            p.n_planes = len(cols["X-shift (pixels)"])
            p.shift_x_3d = cols["X-shift (pixels)"].ctypes.data_as(dp)
            p.shift_y_3d = cols["Y-shift (pixels)"].ctypes.data_as(dp)
            p.theta_3d = cols["Theta (degrees)"].ctypes.data_as(dp)
            p.roi_3d = cols["ROI #"].ctypes.data_as(dp)
            p.int_coeffs_3d = cols["Intensity scaling"].ctypes.data_as(dp)
            p.background_3d = self._bg
            p.axial_offset = float(axialOffset)
            p.pixel_size = float(pixelSize)
            p.axial_magnification = int(axialMagnification)
        self._eng = Engine(p)
        e = self._eng
        self.widthS, self.heightS, self.nPlanes, self.nPlanesM = (e.widthS, e.heightS,
                                                                  e.nPlanes, e.nPlanesM)
        self.widthM, self.heightM = e.wM, e.hM
        self.magnification = int(magnification)
        self.axialMagnification = (self.nPlanesM // self.nPlanes
                                   if calibTablePath3DSRRF is not None else 1)
        self.nFrameForSRRF = int(nFrameForSRRF)
        self.width, self.height = int(width), int(height)

    def _e(self) -> Engine:
        if self._eng is None:
            raise RuntimeError("initialise() first")
        return self._eng

    # -- :583-595
    def loadDriftXYGPUbuffer(self, driftX, driftY):
        e = self._e()
        dx = np.ascontiguousarray(driftX, dtype=np.float32)
        dy = np.ascontiguousarray(driftY, dtype=np.float32)
        e.check(e.L.esrrf_load_drift(e.h, C.f32ptr(dx), C.f32ptr(dy), min(dx.size, dy.size)))

    # -- :636-746
    def calculateSRRF(self, imsRawData, cancel=None) -> bool:
        e = self._e()
        a, dt = self._frames(imsRawData)
        flag = ctypes.c_int(0)
        cancelled = ctypes.c_int(0)
        if cancel is not None and cancel():
            flag.value = 1
        e.check(e.L.esrrf_calculate(e.h, a.ctypes.data_as(ctypes.c_void_p), dt, a.shape[0],
                                    ctypes.byref(flag) if cancel is not None else None,
                                    ctypes.byref(cancelled)))
        return bool(cancelled.value)

    # -- :749-794
    def readSRRFbuffer(self):
        e = self._e()
        out = np.empty((4 * e.nPlanesM, e.hM, e.wM), np.float32)
        e.check(e.L.esrrf_read_result(e.h, C.f32ptr(out)))
        self.imsSRRF = out
        return out

    # -- :797-878
    def readGradientBuffers(self):
        e = self._e()
        if self._eng.params.do_fht_interpolation and self._calib is None:
            out = np.empty((2, e.hM, e.wM), np.float32)
        else:
            k = 3 if self._calib is not None else 2
            out = np.empty((k * e.nPlanes, 2 * e.heightS, 2 * e.widthS), np.float32)
        e.check(e.L.esrrf_read_gradients(e.h, C.f32ptr(out)))
        return out

    # -- :881-905
    def readAlignedPixels(self):
        e = self._e()
        out = np.empty((e.nPlanes, e.heightS, e.widthS), np.float32)
        e.check(e.L.esrrf_read_aligned(e.h, C.f32ptr(out)))
        return out

    # -- :909-935
    def readMPmaps(self):
        e = self._e()
        M = self.magnification
        n = 3 * self.axialMagnification
        out = np.empty((n, M, M), np.float32)
        e.check(e.L.esrrf_read_mp_maps(e.h, C.f32ptr(out)))
        return out

    # -- :939-951
    def release(self):
        if self._eng is not None:
            self._eng.close()
            self._eng = None

    # -- :955-960
    def resetFramePosition(self):
        e = self._e()
        e.check(e.L.esrrf_reset_frame_position(e.h))

    # -- :1009-1015
    def finishQueue(self):
        if self._eng is not None:
            self._eng.check(self._eng.L.esrrf_finish(self._eng.h))

    # -- :963-994
    def reshapeImageStack3D(self, ims):
        """Montage frames -> [n][nPlanes][heightS][widthS] float planes, (p - background) /
        intensity scaling per calibration cell, computed on the device (esrrf_reshape_3d)."""
        e = self._e()
        a, dt = self._frames(ims)
        out = np.empty((a.shape[0], e.nPlanes, e.heightS, e.widthS), np.float32)
        e.check(e.L.esrrf_reshape_3d(e.h, a.ctypes.data_as(ctypes.c_void_p), dt, a.shape[0],
                                     C.f32ptr(out)))
        return out

    # -- :597-633
    def prepareDataFHT_SRRF(self, imsRawData):
        """FHT mode: upload, Fourier-upsample and form the gradients of a load without the
        RGC (esrrf_prepare_fht); readGradientBuffers then returns frame 0's gradients."""
        e = self._e()
        a, dt = self._frames(imsRawData)
        e.check(e.L.esrrf_prepare_fht(e.h, a.ctypes.data_as(ctypes.c_void_p), dt, a.shape[0]))

    @staticmethod
    def _frames(ims):
        a = np.asarray(ims)
        if a.ndim == 2:
            a = a[None]
        if a.dtype == np.uint16:
            return np.ascontiguousarray(a), C.ESRRF_U16
        return np.ascontiguousarray(a, dtype=np.float32), C.ESRRF_F32

    def autotune(self, reps=2):
        """Time the tiled kernel's launch shapes on the current load and keep the fastest
        (esrrf_autotune; no counterpart in the reference)."""
        e = self._e()
        nw, ppw, tl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = ctypes.c_float()
        e.check(e.L.esrrf_autotune(e.h, int(reps), ctypes.byref(nw), ctypes.byref(ppw),
                                   ctypes.byref(tl), ctypes.byref(ms)))
        return dict(nw=nw.value, ppw=ppw.value, taps_in_lds=tl.value, ms=ms.value)

    def launchShape(self):
        """Launch shape of the tiled 2-D kernel in use (esrrf_launch_shape; no counterpart
        in the reference)."""
        e = self._e()
        v = [ctypes.c_int() for _ in range(4)]
        e.check(e.L.esrrf_launch_shape(e.h, *[ctypes.byref(x) for x in v]))
        return dict(nw=v[0].value, ppw=v[1].value, taps_in_lds=v[2].value,
                    frames_per_pass=v[3].value)

    def setLaunchShape(self, nw=None, taps_in_lds=None, frames_per_pass=None):
        """Pin the tiled 2-D kernel's launch shape (esrrf_set_launch_shape; no counterpart in
        the reference); fields left None keep the shape in use."""
        cur = self.launchShape()
        e = self._e()
        e.check(e.L.esrrf_set_launch_shape(
            e.h, int(cur["nw"] if nw is None else nw),
            int(cur["taps_in_lds"] if taps_in_lds is None else taps_in_lds),
            int(cur["frames_per_pass"] if frames_per_pass is None else frames_per_pass)))
        return self.launchShape()

    def frameCounters(self):
        """nCurrentFrame[0..1] (exact bookkeeping hook)."""
        return self._e().counters()
