"""ctypes binding of libesrrf.so (the C ABI declared in include/esrrf.h).

The library must have been built (``python nanoj-esrrf_amd/build_lib.py``); there is no
fallback path: every entry point raises if the HIP library is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libesrrf.so")

ESRRF_OK = 0
ESRRF_U16 = 0
ESRRF_F32 = 1
_ERRNAMES = {1: "ESRRF_ERR_INVALID", 2: "ESRRF_ERR_HIP", 3: "ESRRF_ERR_STATE",
             4: "ESRRF_ERR_NOMEM", 5: "ESRRF_ERR_UNSUPPORTED"}

# Every symbol include/esrrf.h declares.
EXPORTS = [
    "esrrf_device_count", "esrrf_device_name", "esrrf_fastest_device", "esrrf_create",
    "esrrf_destroy", "esrrf_load_drift", "esrrf_reset_frame_position", "esrrf_calculate",
    "esrrf_calculate_device", "esrrf_prepare_fht", "esrrf_reshape_3d",
    "esrrf_read_result", "esrrf_read_result_device",
    "esrrf_read_gradients", "esrrf_read_aligned", "esrrf_read_mp_maps",
    "esrrf_read_frame_counters", "esrrf_get_dims", "esrrf_finish", "esrrf_set_stream",
    "esrrf_set_upload_chunks", "esrrf_host_alloc", "esrrf_host_free",
    "esrrf_profile_enable", "esrrf_profile_read", "esrrf_bic_constants",
    "esrrf_bic_tso_text", "esrrf_fht_constants", "esrrf_active_samples", "esrrf_plan_rolling", "esrrf_plan_load",
    "esrrf_last_error", "esrrf_version", "esrrf_autotune", "esrrf_launch_shape",
    "esrrf_set_launch_shape",
    "esrrf_rgc_kernel",
    "esrrf_sweep_create", "esrrf_sweep_destroy", "esrrf_sweep_set_stream", "esrrf_sweep_run",
    "esrrf_sweep_run_device", "esrrf_sweep_read", "esrrf_sweep_read_device",
    "esrrf_sweep_shape", "esrrf_sweep_last_error",
    "esrrf_rse_create", "esrrf_rse_destroy", "esrrf_rse_load", "esrrf_rse_value_and_grad",
    "esrrf_rse_best_sigma", "esrrf_rse_descend", "esrrf_rse_counters", "esrrf_rse_last_error",
    "esrrf_squirrel_create", "esrrf_squirrel_destroy", "esrrf_squirrel_set_stream",
    "esrrf_squirrel_dims", "esrrf_squirrel_run_device", "esrrf_squirrel_run",
    "esrrf_squirrel_last_error", "esrrf_sweep_squirrel", "esrrf_qnr_map", "esrrf_remove_nans",
    "esrrf_rolling_create", "esrrf_rolling_destroy", "esrrf_rolling_set_stream",
    "esrrf_rolling_begin", "esrrf_rolling_begin_device", "esrrf_rolling_next",
    "esrrf_rolling_next_device", "esrrf_rolling_last_error",
    "esrrf_frc_create", "esrrf_frc_destroy", "esrrf_frc_set_stream", "esrrf_frc_rings",
    "esrrf_frc_run_device", "esrrf_frc_run", "esrrf_frc_last_error",
    "esrrf_frc_run_device_noalloc", "esrrf_sweep_frc",
    "esrrf_drift_create", "esrrf_drift_destroy", "esrrf_drift_set_stream",
    "esrrf_drift_run_device", "esrrf_drift_run", "esrrf_drift_last_error",
]


class EsrrfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int), ("height", ctypes.c_int), ("magnification", ctypes.c_int),
        ("fwhm", ctypes.c_float), ("sensitivity", ctypes.c_float),
        ("n_frames_on_gpu", ctypes.c_int), ("n_frame_for_srrf", ctypes.c_int),
        ("block_length", ctypes.c_int), ("device", ctypes.c_int),
        ("int_weighting", ctypes.c_int), ("do_mp_map_correction", ctypes.c_int),
        ("do_fht_interpolation", ctypes.c_int), ("n_planes", ctypes.c_int),
        ("shift_x_3d", ctypes.POINTER(ctypes.c_double)),
        ("shift_y_3d", ctypes.POINTER(ctypes.c_double)),
        ("theta_3d", ctypes.POINTER(ctypes.c_double)),
        ("roi_3d", ctypes.POINTER(ctypes.c_double)),
        ("int_coeffs_3d", ctypes.POINTER(ctypes.c_double)),
        ("background_3d", ctypes.c_float), ("axial_offset", ctypes.c_float),
        ("pixel_size", ctypes.c_float), ("axial_magnification", ctypes.c_int),
    ]


class SweepParams(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int), ("height", ctypes.c_int),
        ("magnifications", ctypes.POINTER(ctypes.c_int)), ("n_magnifications", ctypes.c_int),
        ("fwhm", ctypes.POINTER(ctypes.c_float)), ("n_fwhm", ctypes.c_int),
        ("sensitivity", ctypes.POINTER(ctypes.c_float)), ("n_sensitivity", ctypes.c_int),
        ("n_frames", ctypes.POINTER(ctypes.c_int)), ("n_n_frames", ctypes.c_int),
        ("int_weighting", ctypes.c_int), ("do_mp_map_correction", ctypes.c_int),
        ("autotune", ctypes.c_int), ("use_graph", ctypes.c_int), ("device", ctypes.c_int),
        ("do_fht_interpolation", ctypes.c_int),
    ]


class RseParams(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int), ("height", ctypes.c_int), ("magnification", ctypes.c_int),
        ("n_frames", ctypes.c_int), ("crop_size", ctypes.c_int), ("int_weighting", ctypes.c_int),
        ("do_mp_map_correction", ctypes.c_int), ("sigma_max", ctypes.c_float),
        ("device", ctypes.c_int),
    ]


class SquirrelParams(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int), ("height", ctypes.c_int), ("magnification", ctypes.c_int),
        ("crop_border", ctypes.c_int), ("crop_size", ctypes.c_int), ("fix_sigma", ctypes.c_int),
        ("fixed_sigma", ctypes.c_float), ("device", ctypes.c_int),
    ]


class SquirrelResult(ctypes.Structure):
    _fields_ = [
        ("sigma", ctypes.c_double), ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
        ("rse", ctypes.c_double), ("rsp", ctypes.c_double), ("overblur", ctypes.c_int),
        ("evaluations", ctypes.c_int),
    ]


class RollingParams(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int), ("height", ctypes.c_int), ("magnification", ctypes.c_int),
        ("fwhm", ctypes.c_float), ("sensitivity", ctypes.c_float),
        ("n_frame_for_srrf", ctypes.c_int), ("frame_gap", ctypes.c_int),
        ("int_weighting", ctypes.c_int), ("do_mp_map_correction", ctypes.c_int),
        ("chunk_frames", ctypes.c_int), ("device", ctypes.c_int),
    ]


_LIB = None


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} "
                           "(run `python nanoj-esrrf_amd/build_lib.py`)")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, fp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    c_int, c_float = ctypes.c_int, ctypes.c_float
    sig = {
        "esrrf_device_count": [ip],
        "esrrf_device_name": [c_int, ctypes.c_char_p, c_int],
        "esrrf_fastest_device": [ip],
        "esrrf_create": [ctypes.POINTER(Params), ctypes.POINTER(vp)],
        "esrrf_destroy": [vp],
        "esrrf_load_drift": [vp, fp, fp, c_int],
        "esrrf_reset_frame_position": [vp],
        "esrrf_calculate": [vp, vp, c_int, c_int, ip, ip],
        "esrrf_calculate_device": [vp, vp, c_int, c_int, ip, ip],
        "esrrf_read_result": [vp, fp],
        "esrrf_read_result_device": [vp, vp],
        "esrrf_read_gradients": [vp, fp],
        "esrrf_read_aligned": [vp, fp],
        "esrrf_prepare_fht": [vp, vp, c_int, c_int],
        "esrrf_reshape_3d": [vp, vp, c_int, c_int, fp],
        "esrrf_read_mp_maps": [vp, fp],
        "esrrf_read_frame_counters": [vp, ip],
        "esrrf_get_dims": [vp, ip, ip, ip, ip, ip, ip],
        "esrrf_finish": [vp],
        "esrrf_set_stream": [vp, vp],
        "esrrf_set_upload_chunks": [vp, c_int],
        "esrrf_host_alloc": [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)],
        "esrrf_host_free": [vp],
        "esrrf_profile_enable": [vp, c_int],
        "esrrf_profile_read": [vp, c_int, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_long)],
        "esrrf_bic_constants": [c_float, fp, fp, fp, fp],
        "esrrf_bic_tso_text": [c_float, ctypes.c_char_p, c_int],
        "esrrf_fht_constants": [c_float, c_int, fp, fp, fp, fp],
        "esrrf_active_samples": [vp, ctypes.POINTER(ctypes.c_longlong)],
        "esrrf_plan_rolling": [c_int, c_int, c_int, c_int, c_int, ip],
        "esrrf_plan_load": [ip, c_int, c_int, ip, ip],
        "esrrf_last_error": [vp],
        "esrrf_version": [],
        "esrrf_autotune": [vp, c_int, ip, ip, ip, fp],
        "esrrf_launch_shape": [vp, ip, ip, ip, ip],
        "esrrf_set_launch_shape": [vp, c_int, c_int, c_int],
        "esrrf_rgc_kernel": [vp, ctypes.c_char_p, c_int],
        "esrrf_sweep_create": [ctypes.POINTER(SweepParams), ctypes.POINTER(vp)],
        "esrrf_sweep_destroy": [vp],
        "esrrf_sweep_set_stream": [vp, vp],
        "esrrf_sweep_run": [vp, vp, c_int, c_int],
        "esrrf_sweep_run_device": [vp, vp, c_int, c_int],
        "esrrf_sweep_read": [vp, c_int, c_int, c_int, c_int, fp],
        "esrrf_sweep_read_device": [vp, c_int, c_int, c_int, c_int, vp],
        "esrrf_sweep_shape": [vp, c_int, c_int, ip, ip, ip, ip, fp],
        "esrrf_sweep_last_error": [vp],
        "esrrf_rse_create": [ctypes.POINTER(RseParams), ctypes.POINTER(vp)],
        "esrrf_rse_destroy": [vp],
        "esrrf_rse_load": [vp, vp, c_int, c_int],
        "esrrf_rse_value_and_grad": [vp, c_float, c_float, c_float, dp, dp],
        "esrrf_rse_best_sigma": [vp, c_float, c_float, dp, dp, ip],
        "esrrf_rse_descend": [vp, fp, c_int, fp, fp, fp, dp],
        "esrrf_rse_counters": [vp, ctypes.POINTER(ctypes.c_longlong),
                               ctypes.POINTER(ctypes.c_longlong)],
        "esrrf_rse_last_error": [vp],
        "esrrf_squirrel_create": [ctypes.POINTER(SquirrelParams), ctypes.POINTER(vp)],
        "esrrf_squirrel_destroy": [vp],
        "esrrf_squirrel_set_stream": [vp, vp],
        "esrrf_squirrel_dims": [vp, ip, ip],
        "esrrf_squirrel_run_device": [vp, vp, vp, ctypes.POINTER(SquirrelResult), vp, vp],
        "esrrf_squirrel_run": [vp, fp, fp, ctypes.POINTER(SquirrelResult), fp, fp],
        "esrrf_squirrel_last_error": [vp],
        "esrrf_sweep_squirrel": [vp, c_int, c_int, c_int, c_float, dp, dp, dp],
        "esrrf_remove_nans": [fp, c_int, c_int, ctypes.c_double],
        "esrrf_qnr_map": [fp, fp, ctypes.c_long, c_int, c_float, c_float, fp, fp,
                          ctypes.POINTER(ctypes.c_long)],
        "esrrf_rolling_create": [ctypes.POINTER(RollingParams), ctypes.POINTER(vp)],
        "esrrf_rolling_destroy": [vp],
        "esrrf_rolling_set_stream": [vp, vp],
        "esrrf_rolling_begin": [vp, vp, c_int, c_int, ip],
        "esrrf_rolling_begin_device": [vp, vp, c_int, c_int, ip],
        "esrrf_rolling_next": [vp, fp, ip],
        "esrrf_rolling_next_device": [vp, vp, ip],
        "esrrf_rolling_last_error": [vp],
        "esrrf_frc_create": [c_int, c_int, c_int, ctypes.POINTER(vp)],
        "esrrf_frc_destroy": [vp],
        "esrrf_frc_set_stream": [vp, vp],
        "esrrf_frc_rings": [vp, ip, ip],
        "esrrf_frc_run_device": [vp, vp, vp, dp, fp],
        "esrrf_frc_run_device_noalloc": [c_int, vp, vp, vp, c_int, c_int, dp, fp, ip],
        "esrrf_frc_run": [vp, fp, fp, dp, fp],
        "esrrf_frc_last_error": [vp],
        "esrrf_drift_create": [c_int, c_int, c_int, ctypes.POINTER(vp)],
        "esrrf_drift_destroy": [vp],
        "esrrf_drift_set_stream": [vp, vp],
        "esrrf_drift_run_device": [vp, vp, c_int, c_int, fp, fp],
        "esrrf_drift_run": [vp, fp, c_int, c_int, fp, fp],
        "esrrf_drift_last_error": [vp],
        "esrrf_sweep_frc": [vp, vp, c_int, dp],
    }
    strings = ("esrrf_last_error", "esrrf_version", "esrrf_sweep_last_error",
               "esrrf_rse_last_error", "esrrf_squirrel_last_error",
               "esrrf_rolling_last_error", "esrrf_frc_last_error", "esrrf_drift_last_error")
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_char_p if name in strings else c_int
    _LIB = L
    return L


def check(rc, handle=None):
    if rc != ESRRF_OK:
        msg = ""
        if handle:
            m = lib().esrrf_last_error(handle)
            msg = m.decode() if m else ""
        raise EsrrfError(rc, msg)


def f32ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def bic_constants(fwhm: float):
    out = [ctypes.c_float() for _ in range(4)]
    check(lib().esrrf_bic_constants(fwhm, *[ctypes.byref(o) for o in out]))
    return dict(zip(("sigma", "radius", "tSS", "tSO"), (o.value for o in out)))


class HostBuffer:
    """Page-locked host memory of `nbytes` from the engine's runtime (esrrf_host_alloc),
    viewable as a numpy array."""

    def __init__(self, nbytes: int):
        self.ptr = ctypes.c_void_p()
        check(lib().esrrf_host_alloc(nbytes, ctypes.byref(self.ptr)))
        self.nbytes = nbytes

    def array(self, dtype, shape) -> np.ndarray:
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        assert n <= self.nbytes
        raw = (ctypes.c_char * n).from_address(self.ptr.value)
        return np.frombuffer(raw, dtype=dtype).reshape(shape)

    def close(self):
        if self.ptr:
            lib().esrrf_host_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bic_tso_text(fwhm: float) -> str:
    """The $TWOSIGpONE$ text LiveSRRF_CL.java:318 bakes into the BIC/3-D kernel."""
    buf = ctypes.create_string_buffer(64)
    check(lib().esrrf_bic_tso_text(fwhm, buf, 64))
    return buf.value.decode()


def fht_constants(fwhm: float, magnification: int):
    out = [ctypes.c_float() for _ in range(4)]
    check(lib().esrrf_fht_constants(fwhm, magnification, *[ctypes.byref(o) for o in out]))
    return dict(zip(("sigma", "radius", "tSS", "tSO"), (o.value for o in out)))


def plan_rolling(n_slices, n_frame_for_srrf_user, frame_gap_user, rolling, max_frames_on_gpu):
    """gui/LiveSRRF_optimised_.java bookkeeping through the C ABI."""
    out = (ctypes.c_int * 6)()
    check(lib().esrrf_plan_rolling(n_slices, n_frame_for_srrf_user, frame_gap_user,
                                   1 if rolling else 0, max_frames_on_gpu, out))
    plan6 = list(out)
    loads = []
    for r in range(1, plan6[2] + 1):
        for l in range(plan6[4]):
            fi, nf = ctypes.c_int(), ctypes.c_int()
            check(lib().esrrf_plan_load(out, r, l, ctypes.byref(fi), ctypes.byref(nf)))
            loads.append((r, l, fi.value, nf.value))
    return dict(nFrameForSRRF=plan6[0], gap=plan6[1], nSRRFframe=plan6[2],
                nFrameOnGPU=plan6[3], nLoads=plan6[4], plan=loads)
