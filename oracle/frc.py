"""Fourier ring correlation of two images -- CPU restatement, test infrastructure only.

The parameter sweep scores configurations by FRC.calculateFireNumber(odd, even,
FIXED_1_OVER_7) (gui/ParametersSweep_.java:437-446), a class of NanoJ-Core that is not in the
reference tree, so its windowing, binning and smoothing are not available (parity
unpinned).  This restatement uses the standard definition (Banterle et al., J. Struct. Biol. 2013;
Nieuwenhuizen et al., Nat. Methods 2013):

* crop both images to the centred S x S square, S = min(width, height);
* multiply by a separable Tukey window (alpha 0.25) and zero-pad to N, the power of two >= S;
* F1, F2 = 2-D DFTs; ring r of frequency (kx, ky) (signed, |k| <= N/2) is the nearest
  integer to sqrt(kx^2 + ky^2), for r < N/2;
* FRC(r) = sum Re(F1 conj F2) / sqrt(sum |F1|^2 sum |F2|^2) over ring r;
* the resolution is N / r*, r* the first crossing of the fixed 1/7 threshold after ring 0
  (linear between the rings around it), NaN when the curve never crosses.
NaN pixels read as 0 (readSRRFbuffer's images).
"""
from __future__ import annotations

import math

import numpy as np

THRESHOLD = 1.0 / 7.0
ALPHA = 0.25


def tukey(n: int, alpha: float = ALPHA) -> np.ndarray:
    if n == 1:
        return np.ones(1)
    x = np.arange(n, dtype=np.float64) / (n - 1)
    w = np.ones(n)
    lo = x < alpha / 2
    hi = x > 1 - alpha / 2
    w[lo] = 0.5 * (1 + np.cos(2 * np.pi / alpha * (x[lo] - alpha / 2)))
    w[hi] = 0.5 * (1 + np.cos(2 * np.pi / alpha * (x[hi] - 1 + alpha / 2)))
    return w


def ring_index(k2: np.ndarray) -> np.ndarray:
    """round(sqrt(k2)) for integer k2 >= 0 (halves up): the largest r with r^2 - r + 1 <= k2."""
    k2 = np.asarray(k2, np.int64)
    r = np.floor(np.sqrt(k2.astype(np.float64))).astype(np.int64)
    r = np.where((r + 1) * (r + 1) - (r + 1) + 1 <= k2, r + 1, r)
    r = np.where((r > 0) & (r * r - r + 1 > k2), r - 1, r)
    return np.where(k2 == 0, 0, r)


def prepare(img, S: int, N: int) -> np.ndarray:
    a = np.asarray(img, np.float64)
    a = np.where(np.isnan(a), 0.0, a)
    H, W = a.shape
    y0, x0 = (H - S) // 2, (W - S) // 2
    w = tukey(S).astype(np.float32).astype(np.float64)   # the engine's float table
    out = np.zeros((N, N))
    out[:S, :S] = a[y0:y0 + S, x0:x0 + S] * w[:, None] * w[None, :]
    return out


def ring_sums(a, b):
    H, W = np.asarray(a).shape
    S = min(H, W)
    N = 1 << max(1, math.ceil(math.log2(S)))
    F1 = np.fft.fft2(prepare(a, S, N))
    F2 = np.fft.fft2(prepare(b, S, N))
    k = np.arange(N)
    k = np.where(k < N // 2, k, k - N)
    r = ring_index(k[:, None] ** 2 + k[None, :] ** 2)
    R = N // 2
    m = r < R
    num = np.bincount(r[m], (F1 * np.conj(F2)).real[m], R)
    d1 = np.bincount(r[m], (np.abs(F1) ** 2)[m], R)
    d2 = np.bincount(r[m], (np.abs(F2) ** 2)[m], R)
    return num, d1, d2, N


def curve_and_resolution(num, d1, d2, N):
    den = np.sqrt(d1 * d2)
    with np.errstate(invalid="ignore", divide="ignore"):
        c = np.where(den > 0, num / np.where(den > 0, den, 1.0), 0.0)
    res = float("nan")
    for r in range(1, len(c)):
        if c[r] < THRESHOLD:
            if c[r - 1] >= THRESHOLD:
                rs = (r - 1) + (c[r - 1] - THRESHOLD) / (c[r - 1] - c[r])
            else:
                rs = float(r)
            res = N / rs
            break
    return c, res


def frc(a, b):
    """(curve, resolution in pixels) of two equally sized images."""
    return curve_and_resolution(*ring_sums(a, b))
