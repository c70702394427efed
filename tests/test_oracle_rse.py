"""Pins of the CPU RSE restatement (oracle/rse.py) by closed forms, brute force and
invariants (CPU only)."""
import math

import numpy as np
import pytest

from oracle import rse as R


def brute_blur(img, sigma, K):
    """Direct 2-D convolution with the product kernel and clamp-to-edge indices."""
    g = R.gauss_taps(sigma, K)
    H, W = img.shape
    out = np.zeros((H, W))
    for y in range(H):
        for x in range(W):
            acc = 0.0
            for j in range(-K, K + 1):
                for i in range(-K, K + 1):
                    acc += g[j + K] * g[i + K] * img[min(max(y + j, 0), H - 1),
                                                     min(max(x + i, 0), W - 1)]
            out[y, x] = acc
    return out


def test_taps_normalised_symmetric_and_derivative():
    for s in (0.4, 1.3, 3.7):
        g = R.gauss_taps(s, 12)
        assert abs(g.sum() - 1.0) < 1e-14 and np.allclose(g, g[::-1])
        h = 1e-6
        fd = (R.gauss_taps(s + h, 12) - R.gauss_taps(s - h, 12)) / (2 * h)
        assert np.allclose(R.gauss_taps(s, 12, True), fd, atol=1e-8)
        assert abs(R.gauss_taps(s, 12, True).sum()) < 1e-14   # normalisation is preserved


def test_separable_blur_equals_bruteforce_with_clamped_edges():
    rng = np.random.default_rng(0)
    img = rng.normal(size=(7, 9))
    assert np.allclose(R.blur(img, 1.1, 3), brute_blur(img, 1.1, 3), atol=1e-12)


def test_blur_keeps_constants_and_ones():
    assert np.allclose(R.blur(np.full((6, 5), 3.25), 2.0, 10), 3.25, atol=1e-12)


def test_blur_sigma_derivative_matches_central_difference():
    rng = np.random.default_rng(1)
    img = rng.normal(size=(10, 8))
    s, K, h = 1.7, 9, 1e-5
    fd = (R.blur(img, s + h, K) - R.blur(img, s - h, K)) / (2 * h)
    assert np.allclose(R.blur_dsigma(img, s, K), fd, atol=1e-7)


def test_alpha_beta_recovers_an_affine_map_and_rse_zero():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(5, 6))
    a, b = R.alpha_beta(x, 2.5 * x - 7.0)
    assert abs(a - 2.5) < 1e-12 and abs(b + 7.0) < 1e-12
    sr = rng.normal(size=(8, 8))
    ref = 3.0 * R.blur(sr, 1.2, 6) + 4.0
    assert R.rse(ref, sr, 1.2, 6) < 1e-12


def test_rse_invariant_to_affine_changes_of_the_sr_image():
    rng = np.random.default_rng(3)
    sr, ref = rng.random((9, 9)), rng.random((9, 9))
    r0 = R.rse(ref, sr, 1.5, 8)
    assert math.isclose(R.rse(ref, -4.0 * sr + 11.0, 1.5, 8), r0, rel_tol=1e-10)


def test_rse_matches_explicit_least_squares():
    rng = np.random.default_rng(4)
    sr, ref = rng.random((7, 8)), rng.random((7, 8))
    x = R.blur(sr, 0.9, 5).ravel()
    A = np.stack([x, np.ones_like(x)], 1)
    coef, res, *_ = np.linalg.lstsq(A, ref.ravel(), rcond=None)
    assert math.isclose(R.rse(ref, sr, 0.9, 5), math.sqrt(res[0] / x.size), rel_tol=1e-10)


@pytest.mark.parametrize("f,lo,hi,xmin", [
    (lambda x: (x - 0.3) ** 2 + 1.0, 0.0, 2.0, 0.3),
    (lambda x: (x - 1.7) ** 4 - 0.5 * (x - 1.7) ** 2, 1.7, 4.0, 1.7 + 0.5),
    (lambda x: abs(x - 0.7) + 0.1 * x, 0.0, 5.0, 0.7),
    (lambda x: math.cosh(x - 3.0), 0.0, 3.0, 3.0),      # minimum on the boundary
])
def test_brent_finds_the_minimum(f, lo, hi, xmin):
    x, fx, n = R.brent_min(f, lo, hi)
    assert abs(x - xmin) < 1e-6 and n < 200
    assert fx <= f(xmin) + 1e-9


def test_pearson_closed_forms():
    rng = np.random.default_rng(5)
    x = rng.normal(size=(6, 7))
    assert math.isclose(R.pearson(x, 3.0 * x + 2.0), 1.0, rel_tol=1e-12)
    assert math.isclose(R.pearson(x, -0.5 * x + 1.0), -1.0, rel_tol=1e-12)
    y = rng.normal(size=(6, 7))
    assert math.isclose(R.pearson(x, y), float(np.corrcoef(x.ravel(), y.ravel())[0, 1]),
                        rel_tol=1e-12)


def test_error_map_of_an_exact_blur_model():
    """ref = alpha blur(SR) + beta exactly at sigma 1.5 (in SR pixels): the sigma search
    lands on 1.5, alpha/beta are recovered, RSE and the error map vanish, RSP = 1."""
    rng = np.random.default_rng(6)
    M, crop = 2, 1
    sr = rng.random((20, 22)) * 100.0
    K = R.default_K(float(np.float32(4.0 / 2.35482) * np.float32(M)))
    ref = 0.8 * R.blur(R.crop(sr, M, crop), 1.5, K) + 12.0
    ref_full = np.zeros_like(sr)
    c = crop * M
    ref_full[c:-c, c:-c] = ref
    out = R.error_map(ref_full, sr, M, crop)
    assert abs(out["sigma"] - 1.5) < 1e-4
    assert abs(out["alpha"] - 0.8) < 1e-4 and abs(out["beta"] - 12.0) < 1e-3
    assert out["rse"] < 1e-3 and out["emap"].max() < 5e-3 and out["rsp"] > 1 - 1e-9


def test_error_map_nan_reads_as_zero_and_fixed_sigma():
    rng = np.random.default_rng(7)
    sr, ref = rng.random((12, 12)), rng.random((12, 12))
    sr_nan = sr.copy()
    sr_nan[3, 4] = np.nan
    sr0 = sr.copy()
    sr0[3, 4] = 0.0
    a = R.error_map(ref, sr_nan, 2, 1, fix_sigma=True, fixed_sigma=0.6)
    b = R.error_map(ref, sr0, 2, 1, fix_sigma=True, fixed_sigma=0.6)
    assert a["sigma"] == b["sigma"] == float(np.float32(np.float32(0.6) * np.float32(2)))
    assert a["rse"] == b["rse"] and np.array_equal(a["emap"], b["emap"])


def test_qnr_logistic_end_points_and_harmonic_mean():
    """L(res_min) = 0.925 and L(res_max) = 0.075 (the logistic's defining points); QnR is the
    harmonic mean of L and RSP, so QnR = RSP where L = RSP; the first maximum is reported."""
    frc = np.array([0.05, 0.2, 0.125, np.nan], np.float32)
    rsp = np.array([0.925, 0.075, 0.9, 0.5], np.float32)
    q, best, idx = R.qnr_map(frc, rsp, auto_min_max=True)
    assert abs(q[0] - 0.925) < 1e-6 and abs(q[1] - 0.075) < 1e-6
    # L(midpoint) = 0.5 by the symmetry of the two defining points
    assert abs(q[2] - 2 * 0.5 * 0.9 / 1.4) < 1e-6
    assert np.isnan(q[3]) and idx == 0 and best == q[0]


def test_remove_nans_closed_forms():
    """Remove NaNs (radius 2: the 21-pixel disc): a NaN in a linear ramp takes the ramp's
    value (its 20 neighbours pair off symmetrically and the two middle ones equal it), a
    constant map stays constant, a NaN block wider than the kernel keeps its centre NaN
    after one pass, and edge pixels extend the map."""
    y, x = np.mgrid[0:9, 0:11]
    ramp = (3.0 * x - 2.0 * y + 5.0).astype(np.float32)
    r = ramp.copy()
    r[4, 5] = np.nan
    np.testing.assert_array_equal(R.remove_nans(r, 2.0), ramp)
    c = np.full((7, 7), 2.5, np.float32)
    c[0, 0] = c[3, 3] = c[6, 2] = np.nan
    np.testing.assert_array_equal(R.remove_nans(c, 2.0), np.full((7, 7), 2.5, np.float32))
    b = np.ones((15, 15), np.float32)
    b[2:13, 2:13] = np.nan
    one = R.remove_nans(b, 2.0)
    assert np.isnan(one[7, 7]) and one[2, 2] == 1.0
    assert not np.isnan(R.remove_nans(R.remove_nans(one))).any()   # 2 cells per pass
    e = np.arange(16, dtype=np.float32).reshape(4, 4)
    e[0, 0] = np.nan
    # at the corner the disc reads replicated edge pixels: rows 0, 0, 0, 1, 2 give the
    # non-NaN values 1 | 1 2 | 1 2 | 4 4 4 5 6 | 8 8 9, median 4
    assert R.remove_nans(e, 2.0)[0, 0] == np.float32(4.0)
