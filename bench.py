"""Benchmark of the MI355X eSRRF engine on the production 2-D workload.

Workload: 1000 frames of 256x256 uint16 (synthetic blinking filaments, 2000 emitters on
40 filaments, background 200), magnification 5, fwhm ("radius") 1.5, sensitivity 2,
1000 frames per SR frame, intensity weighting and macro-pixel correction on, BIC
interpolation -> one 1280x1280 float32 SR image (AVG, VAR, TAC2, INT).

A step = resetFramePosition + calculateSRRF over the 1000 frames already resident in HBM
+ readSRRFbuffer's finalisation (VAR, MP map, MP correction, NaN scrub) into a device
buffer.  Metric: output Mpx/s = 1280*1280*1000 / step time / 1e6.

Prints ONE JSON line.  The RGC kernel's average duration is measured with HIP events on
the launch stream; its algorithmic FLOPs follow the convention 14 per gated RGC sample +
42 per output pixel-frame.  cpu_baseline times the CPU oracle (C restatement, OpenMP) on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "eSRRF output Mpx/s (x5, 1000 frames); sweep configs/s fused vs looped plain"
PEAK_FP32_TFLOPS = 157.3
W, H, NF, M, FWHM, SENS = 256, 256, 1000, 5, 1.5, 2


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=NF, help="frames per SR frame (default 1000)")
    ap.add_argument("--cpu-sample-frames", type=int, default=48)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mode", choices=("bic", "fht"), default="bic",
                    help="bic: the production workload (default); fht: the same stack through "
                         "the Fourier-upsampling mode")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 3 plain 1 GiB device copies (HBM counter calibration)")
    return ap.parse_args()


def copy_calibration():
    """Three device-to-device copies of 1 GiB: the reference reading for rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters (known traffic 1 GiB read + 1 GiB written per copy)."""
    import torch
    n = 256 * 1024 * 1024
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": 4 * n, "copies": 3}))


def cpu_baseline(stack, nframes):
    """Oracle (C restatement of liveSRRF.cl, OpenMP) on the first `nframes` frames."""
    import numpy as np
    from oracle import oracle as O
    threads = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    o = O.OracleLiveSRRF(W, H, M, FWHM, SENS, nframes, nframes, True, True, False,
                         nthreads=threads)
    sub = stack[:nframes].astype(np.float32)
    t0 = time.perf_counter()
    o.calculateSRRF(sub)
    o.readSRRFbuffer()
    dt = time.perf_counter() - t0
    return {"value": round(W * M * H * M * nframes / dt / 1e6, 3), "unit": "Mpx/s",
            "cores": threads, "kind": "port",
            "sample": f"{nframes} of the 1000 frames (256x256, M=5, fwhm 1.5, S=2, BIC, MP on); "
                      f"{dt:.1f} s on {threads} host threads"}


def load_traffic():
    """HBM bytes per RGC launch from the committed rocprofv3 PMC summary (or None): the
    line `hbm_bytes_per_launch: <int>` of profiles/k_bic_rgc_256x256x1000_m5_traffic.md."""
    p = os.path.join(REPO, "profiles", "k_bic_rgc_256x256x1000_m5_traffic.md")
    try:
        with open(p) as f:
            for line in f:
                if line.strip().startswith("hbm_bytes_per_launch:"):
                    return int(line.split(":", 1)[1].strip())
    except (OSError, ValueError):
        return None
    return None


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus above 1 is not "
                                   "supported (multi-GPU sharding is a design only)",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    import __graft_entry__
    pkg = __graft_entry__.load()
    from nanoj_esrrf_amd import synth
    C = pkg._capi

    nf = a.frames
    stack = synth.production_256x256x1000(n_frames=nf)
    assert stack.shape == (nf, H, W) and stack.dtype == np.uint16

    dev = torch.device("cuda:0")
    d_frames = torch.from_numpy(stack.view(np.int16).reshape(-1)).to(dev)
    d_out = torch.empty(4 * W * M * H * M, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    p = C.Params()
    p.width, p.height, p.magnification = W, H, M
    p.fwhm, p.sensitivity = FWHM, SENS
    p.n_frames_on_gpu, p.n_frame_for_srrf, p.block_length, p.device = nf, nf, 20000, 0
    fht = a.mode == "fht"
    p.int_weighting, p.do_mp_map_correction, p.do_fht_interpolation = 1, 1, int(fht)
    eng = pkg.Engine(p)
    L, h = eng.L, eng.h
    eng.check(L.esrrf_set_stream(h, ctypes.c_void_p(stream.cuda_stream)))
    nact = ctypes.c_longlong()
    eng.check(L.esrrf_active_samples(h, ctypes.byref(nact)))
    fptr = ctypes.c_void_p(d_frames.data_ptr())
    optr = ctypes.c_void_p(d_out.data_ptr())

    def step():
        eng.check(L.esrrf_reset_frame_position(h))
        eng.check(L.esrrf_calculate_device(h, fptr, C.ESRRF_U16, nf, None, None))
        eng.check(L.esrrf_read_result_device(h, optr))

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    eng.check(L.esrrf_profile_enable(h, 1))
    t0 = time.perf_counter()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(a.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev0.elapsed_time(ev1)
    rgc_ms, rgc_n = ctypes.c_double(), ctypes.c_long()
    eng.check(L.esrrf_profile_read(h, 0, ctypes.byref(rgc_ms), ctypes.byref(rgc_n)))
    prep_ms, prep_n = ctypes.c_double(), ctypes.c_long()
    eng.check(L.esrrf_profile_read(h, 1, ctypes.byref(prep_ms), ctypes.byref(prep_n)))
    up_ms, up_n = ctypes.c_double(), ctypes.c_long()
    eng.check(L.esrrf_profile_read(h, 2, ctypes.byref(up_ms), ctypes.byref(up_n)))
    eng.check(L.esrrf_profile_enable(h, 0))

    ms_step = ms_total / a.steps
    outpx = W * M * H * M
    value = outpx * nf / (ms_step / 1e3) / 1e6
    rgc_avg = rgc_ms.value / max(rgc_n.value, 1)
    flops = 14.0 * nact.value * nf + 42.0 * outpx * nf
    achieved = flops / (rgc_avg / 1e3) / 1e12

    res = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mpx/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded blinking Bezier filaments, Poisson + read noise, uint16)",
        "config": {"workload": "production 2-D eSRRF: 1000 frames 256x256 uint16 -> 1280x1280, "
                               "M=5, fwhm 1.5, S=2, VAR+AVG+TAC2+INT, MP correction, "
                               + ("FHT upsampling (MP skipped as in the reference)" if fht
                                  else "BIC"),
                   "frames": nf, "width": W, "height": H, "magnification": M,
                   "fwhm": FWHM, "sensitivity": SENS},
        "roofline": {"bound": "valu", "kernel": "k_fht_rgc" if fht else "k_bic_rgc",
                     "achieved": round(achieved, 2), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": None if fht else load_traffic(),
                     "kernel_ms": round(rgc_avg, 3), "launches": rgc_n.value,
                     "algorithmic_flop_per_launch": flops,
                     "gated_samples_per_frame": nact.value,
                     "prep_kernel_ms": round(prep_ms.value / max(prep_n.value, 1), 3)
                     if not fht else None,
                     "upsample_ms_per_load": round(up_ms.value / max(up_n.value, 1), 3)
                     if fht else None},
        "wall_s": round(wall, 3),
    }
    eng.close()
    del d_frames, d_out
    if not a.no_cpu_baseline and not fht:
        res["cpu_baseline"] = cpu_baseline(stack, min(a.cpu_sample_frames, nf))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
