"""SQUIRREL resolution-scaled error (RSE) of an SR image -- CPU restatement, test
infrastructure only.

Follows ErrorMapLiveSRRF.java:
  * crop cropSize*M pixels on every side of both images (:91-103, :61-68);
  * blur the SR image and an image of ones with a Gaussian RSF of sigma (SR pixels) (:137-141);
  * alpha, beta = least squares of ref on (blur(SR), blur(1)) (:201-233);
  * RSE = sqrt(mean((ref - blur(alpha*SR + beta))^2)) (:159-175);
  * unfixed sigma: Brent minimisation of RSE(sigma) on [0, 4/2.35482*M] (:75, :119-127).
ref is the INT channel and SR the AVG (or VAR/TAC2) channel of readSRRFbuffer
(gui/ParametersSweep_.java:473).

``FloatProcessor.blurGaussian`` (ImageJ) is not available here.  Its stand-in, shared with
the HIP engine, is a separable Gaussian g_k = exp(-k^2/(2 sigma^2)) / sum_j exp(-j^2/(2 sigma^2))
over |k| <= K with K = ceil(5 sigma_max) fixed per problem (so RSE is smooth in sigma) and
clamp-to-edge borders (ImageJ extends images by their edge pixels).  With clamp-to-edge the
blurred ones are exactly 1, so beta multiplies a constant image.  Parity with ImageJ's
kernel truncation, edge correction and down-scaling for large sigma is unpinned.

Sums run in float64 (the reference sums in float32).
"""
from __future__ import annotations

import math

import numpy as np


def gauss_taps(sigma: float, K: int, deriv: bool = False) -> np.ndarray:
    k = np.arange(-K, K + 1, dtype=np.float64)
    if sigma <= 0:
        g = (k == 0).astype(np.float64)
        return np.zeros_like(g) if deriv else g
    e = np.exp(-k * k / (2.0 * sigma * sigma))
    Z = e.sum()
    if not deriv:
        return e / Z
    de = e * k * k / sigma ** 3
    return (de * Z - e * de.sum()) / (Z * Z)


def blur1d(img: np.ndarray, taps: np.ndarray, axis: int) -> np.ndarray:
    K = (len(taps) - 1) // 2
    n = img.shape[axis]
    idx = np.clip(np.arange(n)[:, None] + np.arange(-K, K + 1)[None, :], 0, n - 1)
    if axis == 1:
        return np.einsum("yik,k->yi", img[:, idx], taps)
    return np.einsum("ikx,k->ix", img[idx, :], taps)


def blur(img: np.ndarray, sigma: float, K: int) -> np.ndarray:
    g = gauss_taps(sigma, K)
    return blur1d(blur1d(np.asarray(img, np.float64), g, 1), g, 0)


def blur_dsigma(img: np.ndarray, sigma: float, K: int) -> np.ndarray:
    """d/dsigma of blur(img, sigma) (product rule over the two separable passes)."""
    g, dg = gauss_taps(sigma, K), gauss_taps(sigma, K, True)
    x = np.asarray(img, np.float64)
    return blur1d(blur1d(x, g, 0), dg, 1) + blur1d(blur1d(x, dg, 0), g, 1)


def crop(img: np.ndarray, M: int, crop_size: int) -> np.ndarray:
    c = crop_size * M
    return img[c:img.shape[0] - c, c:img.shape[1] - c] if c > 0 else img


def alpha_beta(x: np.ndarray, y: np.ndarray):
    """calculateAlphaBeta (ErrorMapLiveSRRF.java:201-233) with blurred ones == 1."""
    N = float(x.size)
    sxy, sy, sx, sxx = (x * y).sum(), y.sum(), x.sum(), (x * x).sum()
    a = (N * sxy - sy * sx) / (N * sxx - sx * sx)
    b = sy / N - a * (sx / N)
    return a, b


def rse(ref: np.ndarray, sr: np.ndarray, sigma: float, K: int) -> float:
    x = blur(sr, sigma, K)
    y = np.asarray(ref, np.float64)
    a, b = alpha_beta(x, y)
    return math.sqrt(float(((y - a * x - b) ** 2).mean()))


def default_K(sigma_max: float) -> int:
    return max(1, int(math.ceil(5.0 * sigma_max)))


def brent_min(f, lo: float, hi: float, rel=1e-10, abs_=1e-14, max_eval=1000):
    """Brent's derivative-free minimisation (R. P. Brent, Algorithms for Minimization
    without Derivatives, 1973, ch. 5 "localmin"): golden-section steps with parabolic
    interpolation; the routine commons-math3's BrentOptimizer implements.  Returns
    (x, f(x), evaluations)."""
    c = 0.5 * (3.0 - math.sqrt(5.0))
    a, b = lo, hi
    x = w = v = a + c * (b - a)
    fx = fw = fv = f(x)
    n = 1
    d = e = 0.0
    while n < max_eval:
        m = 0.5 * (a + b)
        tol = rel * abs(x) + abs_
        t2 = 2.0 * tol
        if abs(x - m) <= t2 - 0.5 * (b - a):
            break
        p = q = r = 0.0
        if abs(e) > tol:
            r = (x - w) * (fx - fv)
            q = (x - v) * (fx - fw)
            p = (x - v) * q - (x - w) * r
            q = 2.0 * (q - r)
            if q > 0:
                p = -p
            else:
                q = -q
            r = e
            e = d
        if abs(p) < abs(0.5 * q * r) and q * (a - x) < p < q * (b - x):
            d = p / q
            u = x + d
            if u - a < t2 or b - u < t2:
                d = tol if x <= m else -tol
        else:
            e = (b - x) if x < m else (a - x)
            d = c * e
        u = x + (d if abs(d) >= tol else (tol if d > 0 else -tol))
        fu = f(u)
        n += 1
        if fu <= fx:
            if u < x:
                b = x
            else:
                a = x
            v, fv, w, fw, x, fx = w, fw, x, fx, u, fu
        else:
            if u < x:
                a = u
            else:
                b = u
            if fu <= fw or w == x:
                v, fv, w, fw = w, fw, u, fu
            elif fu <= fv or v == x or v == w:
                v, fv = u, fu
    return x, fx, n


def best_sigma(ref, sr, sigma_max: float, K: int):
    """ErrorMapLiveSRRF.optimise with fixSigma false: Brent on [0, sigma_max]."""
    return brent_min(lambda s: rse(ref, sr, s, K), 0.0, sigma_max)
