"""Drift of each frame of a window against its first frame -- CPU restatement, test
infrastructure only (the HIP engine is compared against it; the product never imports it).

Follows XYShiftCalculator.calculateShiftArray (XYShiftCalculator.java:25-47): the reference
frame is the window's first frame (:27), every frame of the window is cross-correlated with
it (:35, NanoJ-Core's CrossCorrelationMap.calculateCrossCorrelationMap(ref, stack, true))
and the shift is read from the correlation peak (:36,
EstimateShiftAndTilt.getShiftFromCrossCorrelationPeak(ccm, MAX_FITTING)), negated into the
sign loadDriftXYGPUbuffer takes (:42-43).  NanoJ-Core is not in the reference tree, so the
map's normalisation and cropping and the peak fit are parity unpinned.  This restatement is
the standard method:

  * both frames minus their means, zero-padded to the next power-of-two square N;
  * CC(t) = sum_x r(x) f(x + t), circular over N, by FFT;
  * the integer peak: the first maximum in row-major order, lags of N/2 and above wrapped to
    negative;
  * per axis, the three-point parabola through the peak and its two neighbours.

The returned drift d is the displacement of a frame's content relative to the reference:
f(x) = r(x - d) gives d.
"""
from __future__ import annotations

import numpy as np


def next_pow2(n: int) -> int:
    N = 1
    while N < n:
        N *= 2
    return N


def padded(img: np.ndarray, N: int) -> np.ndarray:
    a = np.asarray(img, np.float64)
    a = np.where(np.isnan(a), 0.0, a)
    out = np.zeros((N, N))
    out[:a.shape[0], :a.shape[1]] = a - a.mean()
    return out


def cross_correlation(ref: np.ndarray, frame: np.ndarray, N: int) -> np.ndarray:
    """CC[ty][tx] = sum_x r(x) f(x + t), circular over N x N."""
    R = np.fft.fft2(padded(ref, N))
    F = np.fft.fft2(padded(frame, N))
    return np.real(np.fft.ifft2(np.conj(R) * F))


def peak_shift(cc: np.ndarray):
    """(dx, dy) of the first maximum (NaN values never win), refined by a parabola per
    axis.  A map whose maximum is not finite -- a frame holding +-Inf -- gives no shift;
    a parabola through a non-finite neighbour leaves that axis on the integer peak."""
    N = cc.shape[0]
    nn = np.where(np.isnan(cc), -np.inf, cc)
    i = int(np.argmax(nn))
    py, px = divmod(i, N)
    c0 = cc[py, px]
    if not np.isfinite(c0):
        return 0.0, 0.0

    def fit(cm, cp):
        den = cm - 2.0 * c0 + cp
        with np.errstate(invalid="ignore", divide="ignore"):
            f = 0.5 * (cm - cp) / den if den != 0.0 else 0.0
        return f if np.isfinite(f) else 0.0

    dx = fit(cc[py, (px - 1) % N], cc[py, (px + 1) % N])
    dy = fit(cc[(py - 1) % N, px], cc[(py + 1) % N, px])
    lx = px if px < N // 2 else px - N
    ly = py if py < N // 2 else py - N
    return lx + dx, ly + dy


def shift_array(stack: np.ndarray, index_start: int, n: int):
    """shiftX, shiftY of frames index_start .. index_start + n - 1 (0-based) against frame
    index_start (calculateShiftArray with ImageJ's 1-based indexStart minus one)."""
    st = np.asarray(stack)
    h, w = st.shape[1:]
    N = next_pow2(max(w, h))
    ref = st[index_start]
    sx = np.zeros(n)
    sy = np.zeros(n)
    for s in range(n):
        sx[s], sy[s] = peak_shift(cross_correlation(ref, st[index_start + s], N))
    return sx, sy
