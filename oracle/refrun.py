"""ctypes binding of oracle/_ref/libref.so: the reference's OpenCL kernels (compiled from
their own sources by oracle/ref/build_ref.py) run through the AMD OpenCL runtime with
LiveSRRF_CL's enqueue order.  TEST INFRASTRUCTURE ONLY (GPU tests pin the oracle and the
HIP engine against it)."""
from __future__ import annotations

import ctypes
import json
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(_HERE, "_ref")
_LIB = None


def available() -> bool:
    return os.path.exists(os.path.join(REF_DIR, "libref.so"))


def configs():
    with open(os.path.join(REF_DIR, "configs.json")) as f:
        return json.load(f)


def lib():
    global _LIB
    if _LIB is None:
        L = ctypes.CDLL(os.path.join(REF_DIR, "libref.so"))
        fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.ref_run.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int, ctypes.c_int, fp, ip, ctypes.c_int,
                              fp, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, fp,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.ref_run.restype = ctypes.c_int
        L.ref_last_error.restype = ctypes.c_char_p
        _LIB = L
    return _LIB


def st3d_of(theta_deg, shift_x, shift_y):
    """[theta (rad, sign flipped), shiftX, shiftY] as LiveSRRF_CL.java:440-445 uploads."""
    n = len(theta_deg)
    st = np.zeros(3 * n, np.float32)
    for i in range(n):
        st[i] = np.float32(-float(theta_deg[i]) * math.pi / 180.0)
        st[i + n] = np.float32(shift_x[i])
        st[i + 2 * n] = np.float32(shift_y[i])
    return st


def run(name, frames, loads=None, drift=None, st3d=None, do_mp=True, block_length=20000,
        want_gint=False):
    """frames: BIC [n][h][w]; FHT [n][hM][wM] (already upsampled); 3-D [n][nP][h][w]
    (already reshaped).  Returns (out [4][nvox] with NaN -> 0, mp or None, gint or None)."""
    mode_s, w, h, nP, M, fwhm, S, nF, iw = configs()[name]
    mode = {"bic": 0, "fht": 1, "3d": 2}[mode_s]
    fr = np.ascontiguousarray(frames, dtype=np.float32).reshape(-1)
    n = np.asarray(frames).shape[0]
    ld = np.asarray(loads if loads is not None else [n], dtype=np.int32)
    assert int(ld.sum()) == n
    wM, hM = w * M, h * M
    nvox = wM * hM * (nP * M if mode == 2 else 1)
    dr = np.zeros(2 * nF, np.float32) if drift is None else np.ascontiguousarray(drift, np.float32)
    out = np.zeros(4 * nvox, np.float32)
    md = M * M * (M if mode == 2 else 1)
    mp = np.zeros(3 * md, np.float32)
    gint = np.zeros((3 if mode == 2 else 2) * 4 * w * h * nP, np.float32) if want_gint else None
    st = None if st3d is None else np.ascontiguousarray(st3d, np.float32)
    rc = lib().ref_run(os.path.join(REF_DIR, name + ".co").encode(), mode, w, h, nP, M, nF,
                       fr, ld, len(ld), dr,
                       st.ctypes.data if st is not None else None, 1 if do_mp else 0,
                       block_length, out, mp.ctypes.data if do_mp else None,
                       gint.ctypes.data if gint is not None else None, 0)
    if rc != 0:
        raise RuntimeError("reference run failed: " + lib().ref_last_error().decode())
    out[np.isnan(out)] = 0
    return out, (mp if do_mp and mode != 1 else None), gint
