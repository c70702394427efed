"""Build the reference's own OpenCL kernels for gfx950 (test infrastructure only).

Compiles /root/reference/resources/{liveSRRF,liveSRRF_noInterpolation,live3DSRRF}.cl where
they lie, with ROCm's clang as OpenCL C 1.2, once per test configuration: the reference
bakes every constant into the program text by replacing $NAME$ placeholders
(LiveSRRF_CL.java:311-342, :496-509); here the same values -- integers as decimal text,
floats as Java's Float.toString text -- are supplied as -D'$NAME$=text' definitions, so the
preprocessor performs the substitution and the source is compiled unmodified.  Also builds
libref.so (ref_harness.cpp, the LiveSRRF_CL enqueue order on the AMD OpenCL runtime).
Outputs go only to oracle/_ref/.  Nothing from /root/reference is copied.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(os.path.dirname(HERE), "_ref")
REF = "/root/reference/resources"
CLANG = "/opt/rocm/llvm/bin/clang"

# name: (mode, w, h, nP, M, fwhm, S, nFrameForSRRF, intWeighting)
CONFIGS = {
    "bic_plain_64x64_m5": ("bic", 64, 64, 1, 5, 1.5, 1.0, 100, 1),
    "bic_40x24_m2": ("bic", 40, 24, 1, 2, 1.0, 2.0, 7, 1),
    "bic_30x30_m4": ("bic", 30, 30, 1, 4, 2.5, 2.0, 7, 1),
    "bic_27x18_m5_f3": ("bic", 27, 18, 1, 5, 3.0, 3.0, 7, 1),
    "bic_24x20_m3_noiw": ("bic", 24, 20, 1, 3, 1.5, 2.0, 6, 0),
    "bic_24x24_m5_loads": ("bic", 24, 24, 1, 5, 1.5, 2.0, 11, 1),
    "fht_16x16_m3": ("fht", 16, 16, 1, 3, 1.5, 2.0, 4, 1),
    "fht_16x16_m5": ("fht", 16, 16, 1, 5, 1.0, 1.0, 3, 1),
    "d3_16x16x4_m2": ("3d", 16, 16, 4, 2, 1.5, 2.0, 4, 1),
    "d3_12x12x3_m3": ("3d", 12, 12, 3, 3, 1.5, 1.0, 3, 1),
}


def jf(x) -> str:
    """Java Float.toString of a finite float32: the shortest round-trip digits (numpy's
    Dragon4 in unique mode), written as ddd.ddd for 1e-3 <= |x| < 1e7 and as d.dddE<n>
    otherwise, always with a digit after the point."""
    x = np.float32(x)
    if x == 0:
        return "-0.0" if np.signbit(x) else "0.0"
    sci = np.format_float_scientific(x, unique=True, trim="-", exp_digits=1)
    if len(sci.split("e")[0].lstrip("-").replace(".", "")) == 1:
        # one digit suffices: Java picks the closest of the one- and two-digit decimals
        # that round to x (Float.MIN_VALUE prints 1.4E-45)
        two = np.format_float_scientific(x, precision=1, unique=False, exp_digits=1)
        if np.float32(two) == x and abs(Fraction(two) - Fraction(float(x))) < \
                abs(Fraction(sci) - Fraction(float(x))):
            sci = two
    mant, ex = sci.split("e")
    sign = "-" if mant.startswith("-") else ""
    digits = mant.lstrip("-").replace(".", "")
    ex = int(ex)
    if -3 <= ex <= 6:
        if ex >= 0:
            ip = (digits[:ex + 1]).ljust(ex + 1, "0")
            fp = digits[ex + 1:] or "0"
        else:
            ip, fp = "0", "0" * (-ex - 1) + digits
        return f"{sign}{ip}.{fp}"
    return f"{sign}{digits[0]}.{digits[1:] or '0'}E{ex}"


def defines(mode, w, h, nP, M, fwhm, S, nF, iw):
    f32 = np.float32
    sigma = f32(fwhm) / f32(2.354)
    d = {}
    if mode in ("bic", "3d"):
        radius = f32(int(f32(4.0) * sigma)) / f32(2) + f32(1)
        d.update({
            "MAGNIFICATION": str(M), "SENSITIVITY": jf(S), "GXGYMAGNIFICATION": "2",
            "TWOSIGSQUARE": jf(f32(f32(2) * sigma) * sigma),
            # `"" + 2 * sigma + 1` (LiveSRRF_CL.java:318) concatenates strings
            "TWOSIGpONE": jf(f32(2) * sigma) + "1", "RADIUS": jf(radius),
            "WIDTH": str(w), "HEIGHT": str(h), "WH": str(w * h), "WM": str(w * M),
            "HM": str(h * M), "WHM": str(w * h * M * M), "WINT": str(2 * w),
            "HINT": str(2 * h), "WHINT": str(4 * w * h), "VXY_OFFSET": "0.5",
            "VXY_ARRAYSHIFT": "1", "NFRAMEFORSRRF": str(nF), "INTWEIGHTING": str(iw),
        })
        if mode == "3d":
            d.update({"NPLANES": str(nP), "DIMRATIO": "0.0", "WHD": str(w * h * nP),
                      "WHDM": str(w * h * nP * M * M * M)})
    else:
        Mf = f32(M)
        ar = f32(int(f32(2) * sigma)) + f32(1)
        tss = f32(f32(f32(f32(2) * sigma) * Mf) * sigma) * Mf
        d.update({
            "SENSITIVITY": jf(S), "TWOSIGSQUARE": jf(tss),
            "TWOSIGpONE": jf(f32(f32(f32(2) * sigma) + f32(1)) * Mf),
            "RADIUS": jf(Mf * ar), "WIDTH": str(w * M), "HEIGHT": str(h * M),
            "WH": str(w * M * h * M), "VXY_OFFSET": "0.5", "VXY_ARRAYSHIFT": "1",
            "NFRAMEFORSRRF": str(nF), "INTWEIGHTING": str(iw),
        })
    return d


SOURCES = {"bic": "liveSRRF.cl", "fht": "liveSRRF_noInterpolation.cl", "3d": "live3DSRRF.cl"}


def build_one(name, cfg):
    mode = cfg[0]
    co = os.path.join(OUT, name + ".co")
    cmd = [CLANG, "-x", "cl", "-cl-std=CL1.2", "--target=amdgcn-amd-amdhsa", "-mcpu=gfx950",
           "-O3", "--rocm-path=/opt/rocm", "-w"]
    cmd += [f"-D${k}$={v}" for k, v in defines(*cfg).items()]
    cmd += [os.path.join(REF, SOURCES[mode]), "-o", co]
    subprocess.run(cmd, check=True)
    return co


def main():
    if not os.path.isdir(REF):
        print("reference sources not present; oracle/_ref not built", file=sys.stderr)
        return 0
    os.makedirs(OUT, exist_ok=True)
    for name, cfg in CONFIGS.items():
        build_one(name, cfg)
    with open(os.path.join(OUT, "configs.json"), "w") as f:
        json.dump(CONFIGS, f, indent=1)
    subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-I/opt/rocm/include",
                    os.path.join(HERE, "ref_harness.cpp"), "-o", os.path.join(OUT, "libref.so"),
                    "-L/usr/lib/x86_64-linux-gnu", "-lOpenCL"], check=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
