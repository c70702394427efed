"""Drift estimation (XYShiftCalculator.calculateShiftArray) on the device against the
oracle (oracle/drift.py): shifts within 2e-3 pixel of the float64 restatement (the device
correlates in float32), exact integer shifts of periodic frames, batches of frames, and the
vibration correction of the rolling driver (gui/LiveSRRF_optimised_.java:284-305)."""
import numpy as np
import pytest

from oracle import drift as D

pytestmark = pytest.mark.gpu


def fourier_shift(img, dx, dy):
    """img moved by (dx, dy) (periodic band-limited interpolation): f(x) = img(x - d)."""
    h, w = img.shape
    ky = np.fft.fftfreq(h)[:, None]
    kx = np.fft.fftfreq(w)[None, :]
    return np.real(np.fft.ifft2(np.fft.fft2(img) * np.exp(-2j * np.pi * (kx * dx + ky * dy))))


def drifted_stack(n, w, h, seed):
    from nanoj_esrrf_amd import synth
    base = synth.blinking_stack(8, w, h, max(20, w * h // 14), 4, seed=seed).mean(0)
    rng = np.random.default_rng(seed)
    d = np.cumsum(rng.normal(0, 0.4, (n, 2)), axis=0)
    d[0] = 0
    st = np.stack([fourier_shift(base, dx, dy) for dx, dy in d])
    st = st + rng.normal(0, 0.5, st.shape)
    return st.astype(np.float32), d


def run(pkg, st, ref=0):
    calc = pkg.XYShiftCalculator(st)
    sx, sy = calc.calculateShiftArray(1 + ref, st.shape[0] - ref)
    calc.close()
    return sx, sy


@pytest.mark.parametrize("n,w,h", [(12, 64, 64), (9, 40, 24), (70, 32, 32)])
def test_drift_matches_oracle(pkg, n, w, h):
    st, truth = drifted_stack(n, w, h, seed=w + n)
    sx, sy = run(pkg, st)
    ox, oy = D.shift_array(st, 0, n)
    np.testing.assert_allclose(sx, ox, atol=2e-3)
    np.testing.assert_allclose(sy, oy, atol=2e-3)
    if w == h:   # periodic frames: the estimate follows the applied drift, within the
        # three-point parabola's bias on sharp structures
        np.testing.assert_allclose(sx, truth[:, 0], atol=0.2)
        np.testing.assert_allclose(sy, truth[:, 1], atol=0.2)


def test_integer_circular_shifts_exact(pkg):
    rng = np.random.default_rng(3)
    ref = rng.random((32, 32)).astype(np.float32)
    shifts = [(0, 0), (3, -2), (-7, 5), (1, 1), (15, -16)]
    st = np.stack([np.roll(ref, (dy, dx), axis=(0, 1)) for dx, dy in shifts])
    sx, sy = run(pkg, st)   # lags lie in [-16, 16): all of these come back as applied
    np.testing.assert_allclose(sx, [dx for dx, _ in shifts], atol=1e-4)
    np.testing.assert_allclose(sy, [dy for _, dy in shifts], atol=1e-4)


def test_window_start_and_bounds(pkg):
    st, _ = drifted_stack(10, 32, 32, seed=9)
    calc = pkg.XYShiftCalculator(st)
    sx, sy = calc.calculateShiftArray(4, 5)          # frames 4..8 (1-based) against frame 4
    ox, oy = D.shift_array(st, 3, 5)
    np.testing.assert_allclose(sx, ox, atol=2e-3)
    np.testing.assert_allclose(sy, oy, atol=2e-3)
    assert abs(sx[0]) < 1e-4 and abs(sy[0]) < 1e-4
    with pytest.raises(ValueError):
        calc.calculateShiftArray(8, 5)
    calc.close()


def test_vibration_correction_feeds_the_engine(pkg):
    """LiveSRRF_optimised with correctVibration loads, per SR frame, the shifts of its
    window against the window's first frame: the same result as passing those shifts."""
    st, _ = drifted_stack(12, 24, 20, seed=4)
    st = np.clip(st * 10 + 100, 0, None)
    a = pkg.LiveSRRF_optimised(magnification=3, fwhm=1.5, nFrameForSRRF=6,
                               doRollingAnalysis=True, frameGap=3, correctVibration=True)
    got = a.run(st)
    assert a.last_path == "engine"
    calc = pkg.XYShiftCalculator(st)
    b = pkg.LiveSRRF_optimised(magnification=3, fwhm=1.5, nFrameForSRRF=6,
                               doRollingAnalysis=True, frameGap=3)
    exp = b.run(st, drift=lambda r: calc.calculateShiftArray((r - 1) * 3 + 1, 6))
    calc.close()
    for k in got:
        assert np.array_equal(got[k], exp[k])


def test_inf_pixel_frame_gets_no_shift(pkg):
    """A frame holding Inf makes its correlation map NaN; the device keeps the peak index
    in bounds and reports no shift for it, as the oracle does, and the other frames of the
    batch are unaffected."""
    st, _ = drifted_stack(6, 32, 32, seed=11)
    st[3, 5, 7] = np.inf
    sx, sy = run(pkg, st)
    ox, oy = D.shift_array(st, 0, 6)
    assert sx[3] == 0.0 and sy[3] == 0.0 and ox[3] == 0.0 and oy[3] == 0.0
    keep = [0, 1, 2, 4, 5]
    np.testing.assert_allclose(sx[keep], ox[keep], atol=2e-3)
    np.testing.assert_allclose(sy[keep], oy[keep], atol=2e-3)
