"""Pins of the drift oracle (oracle/drift.py) by closed forms: integer circular shifts are
recovered exactly (the autocorrelation of a real image is even, so the parabola adds 0), a
sub-pixel shift of a Gaussian spot is recovered to a few hundredths of a pixel, and the
sign is that of the content's displacement (f(x) = r(x - d) gives +d)."""
import numpy as np
import pytest

from oracle import drift as D


def gaussian(n, cx, cy, s):
    y, x = np.mgrid[0:n, 0:n]
    return np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))


@pytest.mark.parametrize("dx,dy", [(0, 0), (3, -2), (-7, 5), (1, 1)])
def test_integer_circular_shift_exact(dx, dy):
    rng = np.random.default_rng(5)
    ref = rng.random((32, 32))
    frame = np.roll(ref, (dy, dx), axis=(0, 1))    # f(x) = r(x - d), circular
    gx, gy = D.peak_shift(D.cross_correlation(ref, frame, 32))
    assert abs(gx - dx) < 1e-9 and abs(gy - dy) < 1e-9   # exact up to FFT rounding


@pytest.mark.parametrize("dx,dy", [(0.3, -0.6), (-1.25, 0.4), (2.5, 2.1)])
def test_subpixel_gaussian_spot(dx, dy):
    ref = gaussian(64, 30.0, 33.0, 2.5)
    frame = gaussian(64, 30.0 + dx, 33.0 + dy, 2.5)
    gx, gy = D.peak_shift(D.cross_correlation(ref, frame, 64))
    assert abs(gx - dx) < 0.05 and abs(gy - dy) < 0.05


def test_sign_and_padding_of_non_square_frames():
    # 40 x 24 frames pad to N = 64; the spot moves right and up
    ref = gaussian(64, 20.0, 12.0, 2.0)[:24, :40]
    frame = gaussian(64, 23.0, 10.0, 2.0)[:24, :40]
    gx, gy = D.peak_shift(D.cross_correlation(ref, frame, D.next_pow2(40)))
    assert round(gx) == 3 and round(gy) == -2


def test_shift_array_reference_is_the_window_start():
    st = np.stack([gaussian(32, 15.0 + k, 16.0 - k, 2.0) for k in range(4)])
    sx, sy = D.shift_array(np.concatenate([np.zeros((2, 32, 32)), st]), 2, 4)
    np.testing.assert_allclose(sx, [0, 1, 2, 3], atol=0.02)
    np.testing.assert_allclose(sy, [0, -1, -2, -3], atol=0.02)
    assert abs(sx[0]) < 1e-9 and abs(sy[0]) < 1e-9


def test_nan_pixels_read_as_zero():
    ref = gaussian(32, 15.0, 16.0, 2.0)
    frame = np.roll(ref, (0, 2), axis=(0, 1)).copy()
    frame[0, 0] = np.nan
    gx, gy = D.peak_shift(D.cross_correlation(ref, frame, 32))
    assert abs(gx - 2) < 0.01 and abs(gy) < 0.01


def test_inf_pixel_gives_no_shift():
    """A frame holding +-Inf has an infinite mean, so its padded frame and the whole
    correlation map are NaN: no finite maximum, so no shift (and no read through an unset
    peak index)."""
    ref = gaussian(32, 15.0, 16.0, 2.0)
    frame = np.roll(ref, (1, 2), axis=(0, 1)).copy()
    frame[3, 4] = np.inf
    cc = D.cross_correlation(ref, frame, 32)
    assert np.isnan(cc).all()
    assert D.peak_shift(cc) == (0.0, 0.0)
    frame[3, 4] = -np.inf
    assert D.peak_shift(D.cross_correlation(ref, frame, 32)) == (0.0, 0.0)


def test_non_finite_neighbour_keeps_integer_peak():
    cc = np.zeros((8, 8))
    cc[2, 3] = 5.0
    cc[2, 4] = np.nan      # right neighbour along x
    cc[1, 3], cc[3, 3] = 1.0, 3.0
    dx, dy = D.peak_shift(cc)
    assert dx == 3.0
    assert abs(dy - (2 + 0.5 * (1.0 - 3.0) / (1.0 - 10.0 + 3.0))) < 1e-12
