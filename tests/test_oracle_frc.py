"""Pins of the CPU FRC restatement (oracle/frc.py) by closed forms and brute force."""
import math

import numpy as np

from oracle import frc as F


def test_ring_index_is_rounded_radius():
    k2 = np.arange(0, 5000)
    want = np.array([int(math.floor(math.sqrt(v) + 0.5)) for v in k2])
    assert np.array_equal(F.ring_index(k2), want)


def test_ring_sums_match_a_brute_force_dft():
    rng = np.random.default_rng(0)
    a, b = rng.random((7, 8)), rng.random((7, 8))
    num, d1, d2, N = F.ring_sums(a, b)
    assert N == 8
    A, B = F.prepare(a, 7, 8), F.prepare(b, 7, 8)
    n = np.arange(8)
    E = np.exp(-2j * np.pi * np.outer(n, n) / 8)
    FA, FB = E @ A @ E, E @ B @ E
    bn, b1, b2 = np.zeros(4), np.zeros(4), np.zeros(4)
    for ky in range(8):
        for kx in range(8):
            sy, sx = (ky if ky < 4 else ky - 8), (kx if kx < 4 else kx - 8)
            r = int(math.floor(math.sqrt(sx * sx + sy * sy) + 0.5))
            if r < 4:
                bn[r] += (FA[ky, kx] * np.conj(FB[ky, kx])).real
                b1[r] += abs(FA[ky, kx]) ** 2
                b2[r] += abs(FB[ky, kx]) ** 2
    assert np.allclose(num, bn) and np.allclose(d1, b1) and np.allclose(d2, b2)


def test_identical_scaled_and_negated_images():
    rng = np.random.default_rng(1)
    a = rng.random((40, 37))
    c, res = F.frc(a, 3.0 * a)
    assert np.allclose(c[1:], 1.0) and math.isnan(res)
    c, res = F.frc(a, -a)
    assert np.allclose(c[1:], -1.0) and res == 64 / 1.0


def test_band_limited_signal_crosses_at_its_cutoff():
    """Signal band-limited to radius N/8 (or N/16) plus independent noise in each image: the
    FRC is near 1 inside the band and near 0 well outside, so the 1/7 crossing r* sits at
    the cutoff, pushed outwards by the taper's spectral leakage: a Tukey window whose tapers
    span alpha S / 2 pixels spreads a band edge over about 2 / alpha = 8 frequency bins."""
    N = 256
    for frac in (8, 16):
        rng = np.random.default_rng(2)
        k = np.fft.fftfreq(N) * N
        cut = N / frac
        mask = (k[:, None] ** 2 + k[None, :] ** 2) <= cut ** 2
        s = np.fft.ifft2(np.fft.fft2(rng.normal(size=(N, N))) * mask).real
        s *= 2.0 / s.std()
        c, res = F.frc(s + rng.normal(size=(N, N)), s + rng.normal(size=(N, N)))
        rs = N / res
        assert c[int(cut) // 2] > 0.9 and abs(c[int(2 * cut)]) < 0.2
        assert cut <= rs <= cut + 2 / F.ALPHA, (frac, rs)
