"""SQUIRREL error map of (ref, SR) pairs (ErrorMapLiveSRRF.java), the sweep's error mapping
(gui/ParametersSweep_.java:465-530) and the QnR map (gui/GetQnRmap_.java), on the GPU,
against the CPU restatement in oracle/rse.py.

* fixed sigma: RSC and the error map within the parity tolerance (1e-4 relative, 1e-5
  absolute, after division by the oracle's max), RSE and alpha within 1e-5 relative, RSP
  within 1e-6
* Brent sigma: the device search and the oracle's land within 1e-3 relative of each other
  (the objective is flat at its minimum, and the device blurs in float) and their RSE agree
  within 1e-5 relative
* closed form: ref = alpha blur(SR) + beta gives back sigma, alpha, beta, RSE = 0, RSP = 1
"""
import math

import numpy as np
import pytest

from oracle import rse as R
from conftest import normalised_close

pytestmark = pytest.mark.gpu

W, H, M, NF, CROP = 30, 26, 3, 10, 2


@pytest.fixture(scope="module")
def images(pkg):
    """AVG, VAR, TAC2 and INT of one reconstruction by the engine (readSRRFbuffer)."""
    from nanoj_esrrf_amd import synth
    st = synth.blinking_stack(NF, W, H, 60, 4, seed=515)
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(W, H, M, 1.5, 2, NF, NF, 20000, e.allCLdevices[0], True, True, False, None, 0,
                 1)
    e.resetFramePosition()
    e.loadDriftXYGPUbuffer(np.zeros(NF), np.zeros(NF))
    e.calculateSRRF(st)
    out = np.asarray(e.readSRRFbuffer()).reshape(4, H * M, W * M)
    e.release()
    return out


@pytest.mark.parametrize("ch", [0, 1, 2])
def test_fixed_sigma_maps_match_oracle(pkg, images, ch):
    q = pkg.ErrorMapLiveSRRF((H, W), M, fixSigma=True, fixedSigma=0.8, cropBorder=True,
                             cropSize=CROP)
    q.optimise(images[3], images[ch])
    q.calculateErrorMap()
    exp = R.error_map(images[3], images[ch], M, CROP, fix_sigma=True, fixed_sigma=0.8)
    assert q.sigma_linear == exp["sigma"]
    for name, got, want in (("RSC", q.fpSRC, exp["rsc"]), ("error map", q.fpErrorMap,
                                                           exp["emap"])):
        ok, ex, at = normalised_close(got, want)
        assert ok, f"channel {ch} {name}: excess {ex:.3g} at {at}"
    assert math.isclose(q.globalRMSE, exp["rse"], rel_tol=1e-5)
    assert math.isclose(q.alpha, exp["alpha"], rel_tol=1e-5)
    assert abs(q.globalPPMCC - exp["rsp"]) < 1e-6
    q.close()


@pytest.mark.parametrize("ch", [0, 1])
def test_brent_sigma_matches_oracle(pkg, images, ch):
    q = pkg.ErrorMapLiveSRRF((H, W), M, cropSize=CROP)
    q.optimise(images[3], images[ch])
    q.calculateErrorMap()
    exp = R.error_map(images[3], images[ch], M, CROP)
    assert math.isclose(q.sigma_linear, exp["sigma"], rel_tol=1e-3), (q.sigma_linear,
                                                                      exp["sigma"])
    assert math.isclose(q.globalRMSE, exp["rse"], rel_tol=1e-5)
    assert abs(q.globalPPMCC - exp["rsp"]) < 1e-5
    assert q._res.evaluations > 3
    q.close()


def test_exact_blur_model_recovers_parameters(pkg):
    rng = np.random.default_rng(11)
    w, h, m, crop = 24, 20, 2, 1
    sr = (rng.random((h * m, w * m)) * 100.0).astype(np.float32)
    K = R.default_K(float(np.float32(4.0 / 2.35482) * np.float32(m)))
    inner = R.crop(sr.astype(np.float64), m, crop)
    ref = np.zeros_like(sr)
    c = crop * m
    ref[c:-c, c:-c] = 0.8 * R.blur(inner, 1.5, K) + 12.0
    q = pkg.ErrorMapLiveSRRF((h, w), m, cropSize=crop)
    q.optimise(ref, sr)
    q.calculateErrorMap()
    assert abs(q.sigma_linear - 1.5) < 1e-3
    assert abs(q.alpha - 0.8) < 1e-4 and abs(q.beta - 12.0) < 2e-3
    assert q.globalRMSE < 2e-3 and q.globalPPMCC > 1 - 1e-6
    assert float(q.fpErrorMap.max()) < 2e-2
    q.close()


def test_nan_pixels_read_as_zero(pkg, images):
    sr = images[0].copy()
    sr[10, 11] = np.nan
    sr0 = images[0].copy()
    sr0[10, 11] = 0.0
    res = []
    for img in (sr, sr0):
        q = pkg.ErrorMapLiveSRRF((H, W), M, fixSigma=True, fixedSigma=0.5, cropSize=CROP)
        q.optimise(images[3], img)
        q.calculateErrorMap()
        res.append((q.globalRMSE, q.fpErrorMap.copy()))
        q.close()
    assert res[0][0] == res[1][0] and np.array_equal(res[0][1], res[1][1])


def test_sweep_error_maps_match_per_configuration(pkg):
    """Every configuration of a small sweep: the fused sweep's RSE/RSP maps against the
    oracle's error map of that configuration's (INT, channel) pair."""
    from nanoj_esrrf_amd import synth
    w, h, nf = 24, 22, 12
    st = synth.blinking_stack(nf, w, h, 50, 3, seed=99)
    sw = pkg.ParametersSweep(w, h, [2, 3], [1.0, 2.0], [1, 2], [8, 12], autotune=False)
    sw.run(st)
    for ch, name in ((0, "AVG"), (1, "VAR"), (2, "TAC2")):
        maps = sw.error_maps(name, crop_size=2)
        for im in range(2):
            for inf in range(2):
                for isn in range(2):
                    for ifw in range(2):
                        out = sw.read(im, inf, isn, ifw)
                        exp = R.error_map(out[3], out[ch], [2, 3][im], 2)
                        got_rse = maps["rse"][im, inf, isn, ifw]
                        assert math.isclose(got_rse, exp["rse"], rel_tol=1e-4), \
                            (name, im, inf, isn, ifw, got_rse, exp["rse"])
                        assert abs(maps["rsp"][im, inf, isn, ifw] - exp["rsp"]) < 1e-4
    sw.close()


def test_qnr_map_matches_oracle(pkg):
    rng = np.random.default_rng(12)
    frc = rng.uniform(0.04, 0.25, (4, 5)).astype(np.float32)
    frc[1, 2] = np.nan
    rsp = rng.uniform(0.3, 0.99, (4, 5)).astype(np.float32)
    for auto in (True, False):
        q, best, (x, y) = pkg.qnr_map(frc, rsp, auto_min_max=auto, res_min=0.05, res_max=0.2)
        eq, ebest, eidx = R.qnr_map(frc, rsp, auto_min_max=auto, res_min=0.05, res_max=0.2)
        assert np.allclose(q.ravel(), eq, rtol=1e-6, equal_nan=True)
        assert best == pytest.approx(ebest, rel=1e-6) and y * 5 + x == eidx


@pytest.mark.parametrize("w,h", [(100, 90), (64, 64), (130, 257)])
def test_frc_matches_oracle(pkg, w, h):
    from oracle import frc as F
    rng = np.random.default_rng(w + h)
    N = 128
    k = np.fft.fftfreq(N) * N
    base = np.fft.ifft2(np.fft.fft2(rng.normal(size=(N, N))) *
                        ((k[:, None] ** 2 + k[None, :] ** 2) <= 20 ** 2)).real
    sig = np.tile(base, (3, 3))[:h, :w] * 5.0
    a = (sig + rng.normal(size=(h, w))).astype(np.float32)
    b = (sig + rng.normal(size=(h, w))).astype(np.float32)
    a[3, 4] = np.nan
    q = pkg.FRC(w, h)
    res = q.calculateFireNumber(a, b)
    curve, eres = F.frc(a, b)
    assert np.allclose(q.curve, curve, atol=2e-5), float(np.abs(q.curve - curve).max())
    assert math.isclose(res, eres, rel_tol=1e-4), (res, eres)
    assert math.isnan(q.calculateFireNumber(a, a)) and np.allclose(q.curve[1:], 1.0, atol=1e-5)
    q.close()


def test_sweep_frc_maps_match_per_configuration(pkg):
    from nanoj_esrrf_amd import synth
    from oracle import frc as F
    w, h = 24, 22
    st = synth.blinking_stack(24, w, h, 50, 3, seed=98)
    odd, even = pkg.odd_even(st, [8, 12])
    assert np.array_equal(odd[1], st[2]) and np.array_equal(even[0], st[1])
    sw = [pkg.ParametersSweep(w, h, [2, 3], [1.0, 2.0], [1, 2], [8, 12]) for _ in range(2)]
    sw[0].run(odd)
    sw[1].run(even)
    for ch, name in ((0, "AVG"), (2, "TAC2")):
        maps = sw[0].frc_maps(sw[1], name)
        for im in range(2):
            for inf in range(2):
                for isn in range(2):
                    for ifw in range(2):
                        a = sw[0].read(im, inf, isn, ifw)[ch]
                        b = sw[1].read(im, inf, isn, ifw)[ch]
                        _, eres = F.frc(a, b)
                        got = maps[im, inf, isn, ifw]
                        assert (math.isnan(got) and math.isnan(eres)) or \
                            math.isclose(got, eres, rel_tol=1e-3), (name, im, inf, got, eres)
    for s in sw:
        s.close()


def test_remove_nans_matches_oracle(pkg):
    """The library's Remove NaNs pass equals the oracle's on maps with scattered NaNs and
    NaN blocks (odd and even neighbour counts, edges), and qnr_map with removeNaNs runs the
    plugin's passes (RSP three times, FRC twice, RSP once more)."""
    rng = np.random.default_rng(12)
    for shape in [(5, 4), (9, 20), (16, 16)]:
        m = rng.random(shape).astype(np.float32)
        m[rng.random(shape) < 0.3] = np.nan
        m[1:4, 1:3] = np.nan
        np.testing.assert_array_equal(pkg.remove_nans(m, 2.0), R.remove_nans(m, 2.0))
    frc = (0.05 + 0.2 * rng.random((4, 5))).astype(np.float32)
    rsp = rng.random((4, 5)).astype(np.float32)
    frc[1, 2] = np.nan
    rsp[2, 3] = rsp[0, 0] = np.nan
    r = rsp
    for _ in range(4):
        r = R.remove_nans(r, 2.0)
    f = R.remove_nans(R.remove_nans(frc, 2.0), 2.0)
    q, best, idx = pkg.qnr_map(frc, rsp, removeNaNs=True)
    q2, best2, idx2 = pkg.qnr_map(f, r)
    np.testing.assert_array_equal(q, q2)
    assert best == best2 and idx == idx2
