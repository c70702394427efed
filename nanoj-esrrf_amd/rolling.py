"""Rolling analysis of a long stack: the analysis loop of the eSRRF plugin
(gui/LiveSRRF_optimised_.java:238-384) over the C ABI.

``LiveSRRF_optimised.run`` returns the AVG / VAR / TAC2 / INT stacks of every SR frame, or
streams them into a zip of TIFF images as the plugin's "write to disk" option does
(:340-360).  2-D BIC frames without drift go through ``esrrf_rolling`` (one RGC pass per
frame shared by every overlapping SR window); FHT mode and per-window drift loop the engine
per SR frame, as the plugin does (reset, load drift, calculateSRRF per GPU load,
readSRRFbuffer).
"""
from __future__ import annotations

import ctypes
import io
import struct
import zipfile

import numpy as np

from . import _capi as C
from .drift import XYShiftCalculator
from .liveSRRF_CL import LiveSRRF_CL

CHANNELS = ("AVG", "VAR", "TAC2", "INT")


def tiff_bytes(img: np.ndarray) -> bytes:
    """One float32 grayscale image as an uncompressed little-endian TIFF (one strip)."""
    a = np.ascontiguousarray(img, dtype="<f4")
    h, w = a.shape
    data = a.tobytes()
    entries = [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 1, 32), (259, 3, 1, 1),
               (262, 3, 1, 1), (273, 4, 1, 8), (277, 3, 1, 1), (278, 4, 1, h),
               (279, 4, 1, len(data)), (339, 3, 1, 3)]
    ifd = struct.pack("<H", len(entries))
    for tag, typ, cnt, val in entries:
        ifd += struct.pack("<HHI", tag, typ, cnt)
        ifd += struct.pack("<HH", val, 0) if typ == 3 else struct.pack("<I", val)
    ifd += struct.pack("<I", 0)
    return b"II*\x00" + struct.pack("<I", 8 + len(data)) + data + ifd


class Rolling:
    """RAII owner of one ``esrrf_rolling`` (2-D BIC, no drift)."""

    def __init__(self, width, height, magnification, fwhm, sensitivity, n_frame_for_srrf=0,
                 frame_gap=0, int_weighting=True, mp_correction=True, chunk_frames=256,
                 device=-1):
        self.L = C.lib()
        p = C.RollingParams()
        p.width, p.height, p.magnification = int(width), int(height), int(magnification)
        p.fwhm, p.sensitivity = float(fwhm), float(sensitivity)
        p.n_frame_for_srrf, p.frame_gap = int(n_frame_for_srrf), int(frame_gap)
        p.int_weighting = 1 if int_weighting else 0
        p.do_mp_map_correction = 1 if mp_correction else 0
        p.chunk_frames, p.device = int(chunk_frames), int(device)
        h = ctypes.c_void_p()
        C.check(self.L.esrrf_rolling_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h
        self.shape = (4, int(height) * int(magnification), int(width) * int(magnification))
        self._keep = None

    def _check(self, rc):
        if rc != C.ESRRF_OK:
            m = self.L.esrrf_rolling_last_error(self.h)
            raise C.EsrrfError(rc, m.decode() if m else "")

    def begin(self, stack) -> int:
        a = np.asarray(stack)
        dt = C.ESRRF_U16 if a.dtype == np.uint16 else C.ESRRF_F32
        a = np.ascontiguousarray(a if dt == C.ESRRF_U16 else a.astype(np.float32, copy=False))
        self._keep = a     # the engine reads the frames chunk by chunk
        n = ctypes.c_int()
        self._check(self.L.esrrf_rolling_begin(self.h, a.ctypes.data_as(ctypes.c_void_p), dt,
                                               a.shape[0], ctypes.byref(n)))
        return n.value

    def begin_device(self, ptr, dtype, n_slices) -> int:
        n = ctypes.c_int()
        self._check(self.L.esrrf_rolling_begin_device(self.h, ctypes.c_void_p(ptr), dtype,
                                                      int(n_slices), ctypes.byref(n)))
        return n.value

    def next(self):
        out = np.empty(self.shape, np.float32)
        idx = ctypes.c_int()
        self._check(self.L.esrrf_rolling_next(self.h, C.f32ptr(out), ctypes.byref(idx)))
        return idx.value, out

    def next_device(self, ptr) -> int:
        idx = ctypes.c_int()
        self._check(self.L.esrrf_rolling_next_device(self.h, ctypes.c_void_p(ptr),
                                                     ctypes.byref(idx)))
        return idx.value

    def close(self):
        if getattr(self, "h", None):
            self.L.esrrf_rolling_destroy(self.h)
            self.h = None
        self._keep = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LiveSRRF_optimised:
    """The analysis options of the plugin's dialog (gui/LiveSRRF_optimised_.java:540-647)."""

    def __init__(self, magnification=5, fwhm=1.5, sensitivity=1, nFrameForSRRF=0,
                 doRollingAnalysis=False, frameGap=0, intWeighting=True,
                 doMPmapCorrection=True, calculateAVG=True, calculateVAR=False,
                 calculateTAC2=False, getInterpolatedImage=False, doFHTinterpolation=False,
                 chunk_frames=256, correctVibration=False):
        self.magnification = int(magnification)
        self.fwhm, self.sensitivity = float(fwhm), float(sensitivity)
        self.nFrameForSRRF = int(nFrameForSRRF)
        self.doRollingAnalysis, self.frameGap = bool(doRollingAnalysis), int(frameGap)
        self.intWeighting, self.doMPmapCorrection = bool(intWeighting), bool(doMPmapCorrection)
        self.keep = [calculateAVG, calculateVAR, calculateTAC2, getInterpolatedImage]
        self.doFHTinterpolation = bool(doFHTinterpolation)
        self.chunk_frames = int(chunk_frames)
        self.correctVibration = bool(correctVibration)
        self.last_path = None   # "fused" or "engine"

    def plan(self, n_slices):
        return C.plan_rolling(n_slices, self.nFrameForSRRF, self.frameGap,
                              self.doRollingAnalysis, max(n_slices, 1))

    def _frames(self, stack, drift):
        """Yield (r, [4][hM][wM]) for every SR frame, in order."""
        st = np.asarray(stack)
        n_slices, h, w = st.shape
        plan = self.plan(n_slices)
        nF, gap, nSR = plan["nFrameForSRRF"], plan["gap"], plan["nSRRFframe"]
        if self.correctVibration and drift is None:
            # per SR frame, the shifts of its window against its first frame (:284-287)
            calc = XYShiftCalculator(st)
            drift = lambda r: calc.calculateShiftArray((r - 1) * gap + 1, nF)  # noqa: E731
        if not self.doFHTinterpolation and drift is None:
            try:
                r = Rolling(w, h, self.magnification, self.fwhm, self.sensitivity, nF, gap,
                            self.intWeighting, self.doMPmapCorrection, self.chunk_frames)
                n = r.begin(st)
            except C.EsrrfError as e:
                if e.code != 5:        # ESRRF_ERR_UNSUPPORTED: loop the engine instead
                    raise
            else:
                self.last_path = "fused"
                try:
                    for _ in range(n):
                        yield r.next()
                finally:
                    r.close()
                return
        # the plugin's loop: one initialise, per SR frame reset / drift / loads / read (:238-339)
        self.last_path = "engine"
        e = LiveSRRF_CL()
        e.checkDevices()
        e.initialise(w, h, self.magnification, self.fwhm, self.sensitivity, nF, nF, 20000,
                     e.allCLdevices[0], self.intWeighting, self.doMPmapCorrection,
                     self.doFHTinterpolation, None, 0, 1)
        try:
            for r in range(1, nSR + 1):
                e.resetFramePosition()
                dx, dy = (np.zeros(nF, np.float32),) * 2 if drift is None else drift(r)
                e.loadDriftXYGPUbuffer(dx, dy)
                start = (r - 1) * gap          # 0-based indexStartSRRFframe (:281)
                e.calculateSRRF(st[start:start + nF])   # (FHT: upsampled on the device)
                yield r - 1, np.asarray(e.readSRRFbuffer()).reshape(4, h * self.magnification,
                                                                    w * self.magnification)
        finally:
            e.release()

    def run(self, stack, drift=None, writeToDisk=None):
        """Every SR frame of the stack.  drift(r) (1-based SR frame) may return the shift
        arrays (raw pixels, one per frame of the window) that loadDriftXYGPUbuffer takes.
        With writeToDisk (a path), the selected channels of each SR frame are written as
        "eSRRF (<channel>) - frame=<r>.tif" entries of a zip file as they complete
        (:340-360) and the path is returned; otherwise a dict of [nSR][hM][wM] stacks."""
        if writeToDisk is not None:
            with zipfile.ZipFile(writeToDisk, "w", zipfile.ZIP_STORED) as z:
                for idx, out in self._frames(stack, drift):
                    for c, name in enumerate(CHANNELS):
                        if self.keep[c]:
                            z.writestr(f"eSRRF ({name}) - frame={idx + 1}.tif", tiff_bytes(out[c]))
            return writeToDisk
        res = {name: [] for c, name in enumerate(CHANNELS) if self.keep[c]}
        for _, out in self._frames(stack, drift):
            for c, name in enumerate(CHANNELS):
                if self.keep[c]:
                    res[name].append(out[c])
        return {k: np.stack(v) for k, v in res.items()}


def read_tiff(buf: bytes) -> np.ndarray:
    """Read back an image written by tiff_bytes."""
    bio = io.BytesIO(buf)
    assert bio.read(4) == b"II*\x00"
    (ifd,) = struct.unpack("<I", bio.read(4))
    bio.seek(ifd)
    (n,) = struct.unpack("<H", bio.read(2))
    tags = {}
    for _ in range(n):
        tag, typ, cnt = struct.unpack("<HHI", bio.read(8))
        raw = bio.read(4)
        tags[tag] = struct.unpack("<H", raw[:2])[0] if typ == 3 else struct.unpack("<I", raw)[0]
    w, h, off = tags[256], tags[257], tags[273]
    return np.frombuffer(buf[off:off + 4 * w * h], "<f4").reshape(h, w)
