"""Parity at the benchmark's full sizes, through properties that do not depend on size.

* Locality: a magnified pixel's result depends only on raw pixels within a few pixels of
  it (the 2x-grid window of radius 2 r, the 4x4 interpolation support, the 2-point
  gradient), so without the MP correction (a mean over the whole image) the engine's
  result on the full 256x256 x 1000-frame production stack equals, away from the edges of
  a sub-window, the result on that sub-window alone: against the CPU oracle for 2-D BIC
  (|a - b| <= 1e-5 + 1e-4 |b| after division by the channel maximum), and against the
  engine on the sub-montage for the 3-D workload (512 frames of 9 planes of 128x128, M=4).
* MP correction at full size: the corrected channels equal the uncorrected ones divided
  by the MP map computed from them (liveSRRF.cl:464-505), and each sub-pixel position then
  averages to 1.
* The frame-pair and single-frame RGC kernels, and one load against four, give the same
  full-size result bit for bit; so do the FHT mode's two store orders.
"""
import numpy as np
import pytest

from conftest import normalised_close

pytestmark = pytest.mark.gpu

M, FWHM, S = 5, 1.5, 2
MARGIN = 8                 # raw pixels dropped at each sub-window edge (reach <= 7)


@pytest.fixture(scope="module")
def production():
    from nanoj_esrrf_amd import synth
    return synth.production_256x256x1000()


def engine(pkg, st, mp=True, loads=None, fht=False, M_=M, shape=None):
    n, h, w = st.shape
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(w, h, M_, FWHM, S, max(b - a for a, b in (loads or [(0, n)])), n, 20000,
                 e.allCLdevices[0], True, mp, fht, None, 0, 1)
    if shape:
        e.setLaunchShape(**shape)
    e.resetFramePosition()
    e.loadDriftXYGPUbuffer(np.zeros(n), np.zeros(n))
    for a, b in (loads or [(0, n)]):
        e.calculateSRRF(st[a:b])
    out = np.asarray(e.readSRRFbuffer()).copy()
    mpm = np.asarray(e.readMPmaps()).copy() if mp and not fht else None
    e.release()
    return out, mpm


@pytest.fixture(scope="module")
def full_nomp(pkg, production):
    return engine(pkg, production, mp=False)[0]


def test_full_size_bic_matches_oracle_on_a_subwindow(O, production, full_nomp):
    x0, y0, s = 104, 72, 48
    sub = production[:, y0:y0 + s, x0:x0 + s]
    o = O.OracleLiveSRRF(s, s, M, FWHM, S, sub.shape[0], sub.shape[0], True, False, False)
    o.calculateSRRF(sub.astype(np.float32))
    exp = o.readSRRFbuffer()
    a, b = MARGIN * M, (s - MARGIN) * M
    for c, name in enumerate(("AVG", "VAR", "TAC2", "INT")):
        got = full_nomp[c, y0 * M + a:y0 * M + b, x0 * M + a:x0 * M + b]
        ok, ex, at = normalised_close(got, exp[c][a:b, a:b])
        assert ok, f"{name}: excess {ex:.3g} at {at}"


def test_full_size_mp_correction(pkg, production, full_nomp):
    got, mpm = engine(pkg, production, mp=True)
    h, w = production.shape[1:]
    for c in range(3):
        raw = full_nomp[c].astype(np.float64).reshape(h, M, w, M)
        mp = raw.mean(axis=(0, 2))                          # [sy][sx] over macro-pixels
        assert np.allclose(mpm[c], mp, rtol=2e-5), f"MP map {c}"
        want = (raw / mp[None, :, None, :]).reshape(h * M, w * M)
        ok, ex, at = normalised_close(got[c], want)
        assert ok, f"corrected channel {c}: excess {ex:.3g} at {at}"
        means = got[c].astype(np.float64).reshape(h, M, w, M).mean(axis=(0, 2))
        assert np.allclose(means, 1.0, atol=1e-4), f"channel {c} sub-position means"
    assert np.array_equal(got[3], full_nomp[3])          # INT is not corrected (:486-505)


def test_full_size_kernels_and_loads_bit_identical(pkg, production, full_nomp):
    # the default shape runs four frames per pass; one and two frames per pass, and the
    # stack in four loads, give the same bits
    single = engine(pkg, production, mp=False, shape=dict(frames_per_pass=1))[0]
    assert np.array_equal(single, full_nomp)
    pair = engine(pkg, production, mp=False, shape=dict(frames_per_pass=2))[0]
    assert np.array_equal(pair, full_nomp)
    four = engine(pkg, production, mp=False, loads=[(0, 250), (250, 500), (500, 750),
                                                    (750, 1000)])[0]
    assert np.array_equal(four, full_nomp)


def test_full_size_fht_store_orders_bit_identical(pkg, monkeypatch, production):
    st = production[:300]
    monkeypatch.setenv("ESRRF_FHT_SCATTER", "1")
    scatter = engine(pkg, st, fht=True)[0]
    monkeypatch.delenv("ESRRF_FHT_SCATTER")
    tiled = engine(pkg, st, fht=True)[0]
    assert np.array_equal(scatter, tiled)
    assert float(np.abs(tiled[0]).max()) > 0


def test_full_size_3d_locality(pkg):
    from nanoj_esrrf_amd import synth
    from test_gpu_3d import calib_dict, montage, zero_calib
    planes = synth.mfm_9x128x128x512()
    n, nP, h, w = planes.shape
    cd = calib_dict(*zero_calib(nP))

    def run(pl):
        mont = montage(pl, np.arange(nP))
        e = pkg.LiveSRRF_CL()
        e.checkDevices()
        e.initialise(mont.shape[2], mont.shape[1], 4, FWHM, S, n, n, 20000, e.allCLdevices[0],
                     True, False, False, cd, 0, 1)
        e.resetFramePosition()
        e.loadDriftXYGPUbuffer(np.zeros(n), np.zeros(n))
        e.calculateSRRF(mont)
        assert e._e().rgc_kernel() == "k3d_rgc3z"
        out = np.asarray(e.readSRRFbuffer()).reshape(4, nP * 4, pl.shape[2] * 4, pl.shape[3] * 4)
        e.release()
        return out

    full = run(planes)
    x0, y0, s = 40, 56, 32
    sub = run(np.ascontiguousarray(planes[:, :, y0:y0 + s, x0:x0 + s]))
    a, b = MARGIN * 4, (s - MARGIN) * 4
    for c, name in enumerate(("AVG", "VAR", "TAC2", "INT")):
        got = full[c, :, y0 * 4 + a:y0 * 4 + b, x0 * 4 + a:x0 * 4 + b]
        ok, ex, at = normalised_close(got, sub[c, :, a:b, a:b])
        assert ok, f"3-D {name}: excess {ex:.3g} at {at}"
