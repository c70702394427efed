"""Fourier (Hartley) upsampling of one frame -- CPU restatement, test infrastructure only.

Restates ``FourierInterpolation.fhtSpaceInterpolation`` (FourierInterpolation.java:12-73)
and ``mirrorPaddingEvenSquare`` (FourierInterpolation.java:118-172).

The forward/inverse Hartley transforms come from ``nanoj.core2.NanoJFHT``, which is not
vendored with the reference (no source, no version pin).  They are restated here from the
mathematics of the discrete Hartley transform:

* forward:  H(k) = sum_n x(n) cas(2 pi k.n / N) = Re F(k) - Im F(k), F = 2-D DFT (e^{-i}),
  i.e. the true 2-D DHT (the row-column transform plus Bracewell's fix-up, the convention
  the reference's own 3-D FHT shows at fht.cl:497-537);
* inverse:  x(n) = 1/N^2 sum_k H(k) cas(2 pi k.n / N)  (normalised inverse).

With that normalisation the upsampled frame carries intensity scale 1/M^2 (the reference's
3-D variant multiplies the spectrum by M^2 to undo exactly this, FHT3D_Utilities.java:52).
The scale is "parity unpinned": CGLH is scale invariant, AVG/INT scale by c and VAR/TAC2 by
c^2, so images normalised by their maximum are unaffected.

Pins (tests/test_oracle_fht.py): an O(N^4) brute-force cas-sum DHT on N=8, forward-inverse
identity, reproduction of band-limited inputs at every M-th output pixel, and mirror padding
being the identity for power-of-two squares.
"""
from __future__ import annotations

import math

import numpy as np


def _java_log2_ceil(n: int) -> int:
    # (int) Math.ceil(Math.log(n)/Math.log(2))   FourierInterpolation.java:125,128
    return int(math.ceil(math.log(n) / math.log(2)))


def mirror_padding_even_square(img: np.ndarray) -> np.ndarray:
    """FourierInterpolation.java:118-172 (odd mirror-tile factor, centre crop to 2^k square)."""
    h, w = img.shape
    bit_w = int(math.pow(2, _java_log2_ceil(w)))
    bit_h = int(math.pow(2, _java_log2_ceil(h)))
    out_dim = max(bit_h, bit_w)
    scale_w = int(math.ceil(out_dim / w))
    if scale_w % 2 == 0:
        scale_w += 1
    scale_h = int(math.ceil(out_dim / h))
    if scale_h % 2 == 0:
        scale_h += 1
    padded = np.empty((scale_h * h, scale_w * w), dtype=np.float32)
    for sh in range(scale_h):
        flip_v = ((sh - (scale_h - 1) // 2) % 2) != 0
        for sw in range(scale_w):
            flip_h = ((sw - (scale_w - 1) // 2) % 2) != 0
            tile = img
            if flip_h:
                tile = tile[:, ::-1]
            if flip_v:
                tile = tile[::-1, :]
            padded[sh * h:(sh + 1) * h, sw * w:(sw + 1) * w] = tile
    x0 = (scale_w * w - out_dim) // 2
    y0 = (scale_h * h - out_dim) // 2
    return padded[y0:y0 + out_dim, x0:x0 + out_dim].copy()


def dht2(x: np.ndarray) -> np.ndarray:
    """Unnormalised 2-D DHT, H = Re F - Im F (float64)."""
    F = np.fft.fft2(np.asarray(x, dtype=np.float64))
    return F.real - F.imag


def idht2(H: np.ndarray) -> np.ndarray:
    """Normalised inverse 2-D DHT: the DHT is its own inverse up to 1/N^2."""
    n = H.shape[0] * H.shape[1]
    return dht2(H) / n


def zero_insert(H: np.ndarray, M: int) -> np.ndarray:
    """Spectrum insertion of FourierInterpolation.java:33-42 (Nyquist to the negative side)."""
    h, w = H.shape
    out = np.zeros((h * M, w * M), dtype=np.float64)
    p = np.arange(w)
    q = np.arange(h)
    pi = np.where(p < w // 2, p, p + (M - 1) * w)
    qi = np.where(q < h // 2, q, q + (M - 1) * h)
    out[np.ix_(qi, pi)] = H
    return out


def fht_space_interpolation(img: np.ndarray, M: int, mirror: bool = True) -> np.ndarray:
    """FourierInterpolation.java:12-58 for one frame; returns float32 (M*h, M*w)."""
    img = np.asarray(img, dtype=np.float32)
    src = mirror_padding_even_square(img) if mirror else img
    H = dht2(src).astype(np.float32).astype(np.float64)
    Hi = zero_insert(H, M)
    up = idht2(Hi)
    if mirror:
        x0 = (Hi.shape[1] - M * img.shape[1]) // 2
        y0 = (Hi.shape[0] - M * img.shape[0]) // 2
        up = up[y0:y0 + M * img.shape[0], x0:x0 + M * img.shape[1]]
    return up.astype(np.float32)


def fht_stack(stack: np.ndarray, M: int, mirror: bool = True) -> np.ndarray:
    """FourierInterpolation.java:60-73: frame-by-frame upsampling of a stack."""
    return np.stack([fht_space_interpolation(f, M, mirror) for f in stack]).astype(np.float32)


def dht2_bruteforce(x: np.ndarray) -> np.ndarray:
    """O(N^4) cas-sum definition of the 2-D DHT (pin for dht2)."""
    x = np.asarray(x, dtype=np.float64)
    h, w = x.shape
    H = np.zeros((h, w))
    ys, xs = np.mgrid[0:h, 0:w]
    for ky in range(h):
        for kx in range(w):
            ang = 2.0 * np.pi * (kx * xs / w + ky * ys / h)
            H[ky, kx] = np.sum(x * (np.cos(ang) + np.sin(ang)))
    return H
