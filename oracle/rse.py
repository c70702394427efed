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


def finite_or_zero(img) -> np.ndarray:
    """readSRRFbuffer hands over images with NaN -> 0 (LiveSRRF_CL.java:776-781)."""
    a = np.asarray(img, np.float64)
    return np.where(np.isnan(a), 0.0, a)


def pearson(x: np.ndarray, y: np.ndarray) -> float:
    """Pearson correlation (globalPPMCC, ErrorMapLiveSRRF.java:176; NanoJ-Core's
    ArrayMath.calculatePPMCC(.., true) is not available here: its flag is taken to mean the
    mean-centred coefficient, parity unpinned)."""
    x = np.asarray(x, np.float64).ravel()
    y = np.asarray(y, np.float64).ravel()
    xc, yc = x - x.mean(), y - y.mean()
    den = math.sqrt(float((xc * xc).sum() * (yc * yc).sum()))
    return float((xc * yc).sum() / den) if den > 0 else 0.0


def error_map(ref_full, sr_full, M: int, crop_size: int, crop_border: bool = True,
              fix_sigma: bool = False, fixed_sigma: float = 0.0, K: int | None = None):
    """ErrorMapLiveSRRF.optimise + calculateErrorMap (:81-190) on full-size SR-space images.
    Returns dict(sigma, alpha, beta, rse, rsp, emap, rsc, overblur)."""
    ref = crop(finite_or_zero(ref_full), M, crop_size if crop_border else 0)
    sr = crop(finite_or_zero(sr_full), M, crop_size if crop_border else 0)
    sigma_max = float(np.float32(4.0 / 2.35482) * np.float32(M))           # :75
    fixed = float(np.float32(fixed_sigma) * np.float32(M))                 # :53
    if K is None:
        K = default_K(max(sigma_max, fixed if fix_sigma else 0.0))
    if fix_sigma:
        sigma = fixed
    else:
        sigma, _, _ = best_sigma(ref, sr, sigma_max, K)                    # :119-127
    sigma = float(np.float32(sigma))                                       # sigma_linear
    x = blur(sr, sigma, K)
    a, b = alpha_beta(x, ref)                                              # :145-152
    a32, b32 = np.float32(a), np.float32(b)
    rsc = (a32 * x.astype(np.float32) + b32).astype(np.float64)            # :160-168
    d = rsc - ref
    return dict(sigma=sigma, alpha=float(a32), beta=float(b32),
                rse=math.sqrt(float((d * d).mean())),                      # :175
                rsp=pearson(rsc, ref),                                     # :176
                emap=np.abs(ref - rsc),                                    # :180-188
                rsc=rsc, overblur=abs(sigma - sigma_max) < 1e-4)           # :135


def qnr_map(frc, rsp, auto_min_max=True, res_min=0.05, res_max=0.2):
    """gui/GetQnRmap_.java: logistic conversion of the FRC resolution map through
    (res_min, 0.925) and (res_max, 0.075) (:179-203), then QnR = 2 L RSP / (L + RSP)
    (:227-234), first maximum (:244-258).  Returns (qnr, best, index)."""
    frc = np.asarray(frc, np.float32).ravel()
    rsp = np.asarray(rsp, np.float32).ravel()
    if auto_min_max:                                                        # :86-90, :290-305
        f = frc[np.isfinite(frc)]
        res_min, res_max = float(f.min()), float(f.max())
    M1, M2 = 0.075, 0.925
    A1, A2 = math.log((1 - M1) / M1), math.log((1 - M2) / M2)
    x0 = (A2 * float(res_max) - A1 * float(res_min)) / (A2 - A1)
    k = 1 / (x0 - float(res_max)) * A1
    L = (1.0 / (np.exp(-k * (frc.astype(np.float64) - x0)) + 1.0)).astype(np.float32)
    q = (np.float32(2) * L * rsp) / (L + rsp)
    q = q.astype(np.float32)
    best, idx = -np.float32(3.4028235e38), 0
    for i, v in enumerate(q):
        if v > best:
            best, idx = v, i
    return q, float(best), idx


def remove_nans(img, radius=2.0):
    """One pass of ImageJ's "Remove NaNs..." as GetQnRmap_ runs it (gui/GetQnRmap_.java:69-77):
    each NaN pixel becomes the median of the non-NaN pixels of the circular kernel
    dx^2 + dy^2 <= r^2 + 1, edge pixels extending the map, reading the map before the pass.
    ImageJ (RankFilters) is not available: the kernel shape and the even-count median (mean
    of the two middle values) are parity unpinned."""
    a = np.asarray(img, np.float32)
    h, w = a.shape
    r2 = radius * radius + 1.0
    kr = int(math.sqrt(r2 + 1e-10))
    out = a.copy()
    for y in range(h):
        for x in range(w):
            if not np.isnan(a[y, x]):
                continue
            vals = []
            for dy in range(-kr, kr + 1):
                dxm = int(math.sqrt(r2 - dy * dy + 1e-10))
                yy = min(max(y + dy, 0), h - 1)
                for dx in range(-dxm, dxm + 1):
                    v = a[yy, min(max(x + dx, 0), w - 1)]
                    if not np.isnan(v):
                        vals.append(float(v))
            if vals:
                vals.sort()
                n = len(vals)
                out[y, x] = vals[n // 2] if n % 2 else np.float32(0.5 * (vals[n // 2 - 1] + vals[n // 2]))
    return out
