"""FHT-mode (Fourier upsampling + liveSRRF_noInterpolation.cl) parity of the HIP engine
against the CPU oracle (oracle/fht.py for the upsampling, oracle/esrrf_oracle.c for the
gradients and the RGC).

Tolerance: per channel, both images divided by the oracle image's max |value|,
|a - b| <= 1e-5 + 1e-4 |b|.  Frame counters: exact.
"""
import numpy as np
import pytest

from conftest import normalised_close

pytestmark = pytest.mark.gpu


def stack(n, w, h, seed):
    from nanoj_esrrf_amd import synth
    return synth.blinking_stack(n, w, h, max(20, w * h // 14), 4, seed=seed)


def run_engine(pkg, st, M, fwhm, S, intW=True, drift=None, loads=None, nOnGPU=None):
    n = st.shape[0]
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(st.shape[2], st.shape[1], M, fwhm, S, nOnGPU or n, n, 20000,
                 e.allCLdevices[0], intW, True, True, None, 0, 1)
    e.resetFramePosition()
    dx, dy = (np.zeros(n), np.zeros(n)) if drift is None else drift
    e.loadDriftXYGPUbuffer(dx, dy)
    grads = None
    for a, b in (loads or [(0, n)]):
        e.calculateSRRF(st[a:b])
        if grads is None:
            grads = e.readGradientBuffers()
    out = e.readSRRFbuffer()
    counters = e.frameCounters()
    e.release()
    return out, counters, grads


def run_oracle(O, st, M, fwhm, S, intW=True, drift=None, loads=None, nOnGPU=None):
    n = st.shape[0]
    o = O.OracleLiveSRRF(st.shape[2], st.shape[1], M, fwhm, S, nOnGPU or n, n, intW, True, True)
    if drift is not None:
        o.loadDriftXYGPUbuffer(*drift)
    grads = None
    for a, b in (loads or [(0, n)]):
        o.calculateSRRF(st[a:b].astype(np.float32))
        if grads is None:
            grads = o.readGradientBuffers()
    return o.readSRRFbuffer(), list(o.counters), grads


def assert_parity(got, exp, what):
    for c, name in enumerate(("AVG", "VAR", "TAC2", "INT")):
        ok, ex, idx = normalised_close(got[c], exp[c])
        assert ok, f"{what}: {name} differs (excess {ex:.3g} at {idx})"


@pytest.mark.parametrize("M,fwhm,S,w,h", [
    (2, 1.5, 1, 16, 16),      # power-of-two square: mirror padding is the identity
    (3, 1.5, 2, 21, 17),      # odd sizes: mirror tiles 3 x 3 then centre crop to 32
    (4, 2.0, 1, 30, 24),
    (5, 1.5, 2, 32, 20),
    (5, 1.2, 1.7, 24, 24),    # non-integer sensitivity (powf)
    (1, 1.5, 1, 20, 20),      # M = 1: no upsampling
])
def test_fht_configs_parity(pkg, O, M, fwhm, S, w, h):
    st = stack(6, w, h, seed=300 + 11 * M + w)
    got, cnt, g = run_engine(pkg, st, M, fwhm, S)
    exp, ocnt, og = run_oracle(O, st, M, fwhm, S)
    assert cnt == ocnt == [6, 6]
    assert_parity(got.reshape(4, h * M, w * M), exp.reshape(4, h * M, w * M),
                  f"FHT M{M} fwhm{fwhm} S{S} {w}x{h}")
    # RobX gradients of the upsampled first frame (linear in the upsampled pixels)
    g = np.asarray(g).reshape(2, -1)
    og = np.asarray(og).reshape(2, -1)
    for c in range(2):
        scale = np.abs(og[c]).max()
        assert np.abs(g[c] - og[c]).max() <= 2e-5 * scale, f"gradient {c} differs"


@pytest.mark.parametrize("M,w,h,loads", [
    (5, 32, 20, [(0, 5), (5, 9)]), (3, 21, 17, [(0, 7)]), (4, 30, 24, [(0, 1), (1, 4)]),
])
def test_fht_batched_rgc_matches_single_kernel(pkg, monkeypatch, M, w, h, loads):
    """The batched FHT RGC kernel (two pixels and two frames per lane, the default)
    reproduces the one-pixel one-frame kernel bit for bit: odd tails, one-frame loads,
    continuation across loads and ragged tile edges included."""
    n = loads[-1][1]
    st = stack(n, w, h, seed=60 + M)
    res = {}
    for pair in ("0", "1"):
        monkeypatch.setenv("ESRRF_FHT_PAIR", pair)
        e = pkg.LiveSRRF_CL()
        e.checkDevices()
        e.initialise(w, h, M, 1.5, 2, max(b - a for a, b in loads), n, 20000,
                     e.allCLdevices[0], True, True, True, None, 0, 1)
        e.resetFramePosition()
        e.loadDriftXYGPUbuffer(np.zeros(n), np.zeros(n))
        for a, b in loads:
            e.calculateSRRF(st[a:b])
        res[pair] = e.readSRRFbuffer()
        e.release()
    assert np.array_equal(res["0"], res["1"])


@pytest.mark.parametrize("M,w,h", [(5, 32, 20), (3, 21, 17), (2, 16, 16), (4, 40, 33)])
def test_fht_row_tiled_last_pass_matches_scatter(pkg, monkeypatch, M, w, h):
    """The last inverse pass with row-tiled stores (the default) writes exactly what the
    pixel-by-pixel kernel writes: the same column transforms, only the store order differs
    (crops of padded non-power-of-two frames included)."""
    st = stack(6, w, h, seed=80 + M)
    res, grads = {}, {}
    for sc in ("1", "0"):
        monkeypatch.setenv("ESRRF_FHT_SCATTER", sc)
        e = pkg.LiveSRRF_CL()
        e.checkDevices()
        e.initialise(w, h, M, 1.5, 2, 6, 6, 20000, e.allCLdevices[0], True, True, True, None,
                     0, 1)
        e.resetFramePosition()
        e.loadDriftXYGPUbuffer(np.zeros(6), np.zeros(6))
        e.calculateSRRF(st)
        grads[sc] = np.asarray(e.readGradientBuffers()).copy()
        res[sc] = np.asarray(e.readSRRFbuffer()).copy()
        e.release()
    assert np.array_equal(grads["0"], grads["1"])
    assert np.array_equal(res["0"], res["1"])


def test_fht_no_intensity_weighting(pkg, O):
    st = stack(5, 20, 20, seed=17)
    got, _, _ = run_engine(pkg, st, 3, 1.5, 1, intW=False)
    exp, _, _ = run_oracle(O, st, 3, 1.5, 1, intW=False)
    assert_parity(got.reshape(4, 60, 60), exp.reshape(4, 60, 60), "FHT intW off")


def test_fht_uneven_loads_and_counters(pkg, O):
    st = stack(11, 16, 16, seed=23)
    loads = [(0, 4), (4, 8), (8, 11)]
    got, cnt, _ = run_engine(pkg, st, 4, 1.5, 1, loads=loads, nOnGPU=4)
    exp, ocnt, _ = run_oracle(O, st, 4, 1.5, 1, loads=loads, nOnGPU=4)
    assert cnt == ocnt == [11, 3]
    assert_parity(got.reshape(4, 64, 64), exp.reshape(4, 64, 64), "FHT loads 4+4+3")


def test_fht_drift_generic_path(pkg, O):
    st = stack(6, 16, 16, seed=31)
    rng = np.random.default_rng(3)
    drift = (rng.uniform(-0.3, 0.3, 6).astype(np.float32),
             rng.uniform(-0.3, 0.3, 6).astype(np.float32))
    drift[0][2] = drift[1][2] = 0.0          # one unshifted frame inside the series
    got, _, _ = run_engine(pkg, st, 3, 1.5, 1, drift=drift)
    exp, _, _ = run_oracle(O, st, 3, 1.5, 1, drift=drift)
    assert_parity(got.reshape(4, 48, 48), exp.reshape(4, 48, 48), "FHT drift")


def test_fht_mp_maps_are_absent(pkg):
    from nanoj_esrrf_amd import _capi
    st = stack(3, 16, 16, seed=1)
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(16, 16, 2, 1.5, 1, 3, 3, 20000, e.allCLdevices[0], True, True, True, None, 0, 1)
    e.resetFramePosition()
    e.loadDriftXYGPUbuffer(np.zeros(3), np.zeros(3))
    e.calculateSRRF(st)
    with pytest.raises(_capi.EsrrfError):
        e.readMPmaps()
    e.release()


def test_fht_upsampling_of_band_limited_frames(pkg):
    """A frame that is a sum of DFT modes of the 32 x 32 grid (no Nyquist term) is upsampled
    exactly: every M-th magnified pixel reproduces the input, and the magnified frame's
    RobX gradient equals the closed form of the continuous cosine field (up to 1/M^2)."""
    w = h = 32
    M = 3
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    f = 5.0 + np.cos(2 * np.pi * (3 * x + 2 * y) / 32) + 0.5 * np.sin(2 * np.pi * (x - 5 * y) / 32)
    st = np.repeat(f[None].astype(np.float32), 3, axis=0)
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(w, h, M, 1.5, 1, 3, 3, 20000, e.allCLdevices[0], True, False, True, None, 0, 1)
    e.resetFramePosition()
    e.loadDriftXYGPUbuffer(np.zeros(3), np.zeros(3))
    e.calculateSRRF(st)
    out = e.readSRRFbuffer().reshape(4, h * M, w * M)
    e.release()
    # INT channel = mean of the (scale 1/M^2) upsampled frames
    up = out[3] * M * M
    assert np.abs(up[::M, ::M] - f).max() < 1e-4 * np.abs(f).max()
    ym, xm = np.mgrid[0:h * M, 0:w * M].astype(np.float64) / M
    fm = 5.0 + np.cos(2 * np.pi * (3 * xm + 2 * ym) / 32) + \
        0.5 * np.sin(2 * np.pi * (xm - 5 * ym) / 32)
    assert np.abs(up - fm).max() < 2e-4 * np.abs(fm).max()


@pytest.mark.parametrize("var", ["ESRRF_FHT_XCD", "ESRRF_FHT_PREFETCH"])
@pytest.mark.parametrize("M,w,h", [(5, 32, 20), (3, 64, 48), (2, 16, 16)])
def test_fht_tile_order_and_prefetch_do_not_change_results(pkg, monkeypatch, var, M, w, h):
    """The XCD-grouped tile order of the column passes and the register prefetch of the last
    pass only reorder work: with either switched off the upsampled gradients and the SR
    image are the same bit for bit."""
    st = stack(12, w, h, seed=90 + M)
    res, grads = {}, {}
    for on in ("1", "0"):
        monkeypatch.setenv(var, on)
        e = pkg.LiveSRRF_CL()
        e.checkDevices()
        e.initialise(w, h, M, 1.5, 2, 12, 12, 20000, e.allCLdevices[0], True, True, True, None,
                     0, 1)
        e.resetFramePosition()
        e.loadDriftXYGPUbuffer(np.zeros(12), np.zeros(12))
        e.calculateSRRF(st)
        grads[on] = np.asarray(e.readGradientBuffers()).copy()
        res[on] = np.asarray(e.readSRRFbuffer()).copy()
        e.release()
    assert np.array_equal(grads["0"], grads["1"])
    assert np.array_equal(res["0"], res["1"])


@pytest.mark.parametrize("M,w,h", [(3, 20, 17), (1, 16, 12)])
def test_fht_prepare_data_gives_the_load_gradients(pkg, O, M, w, h):
    """prepareDataFHT_SRRF (LiveSRRF_CL.java:597-633): upload, upsampling and gradients of a
    load without the RGC.  The gradients equal those of calculateSRRF on the same frames,
    the oracle's within the parity tolerance, the load counter is reset and the SR frame's
    counter and result are untouched."""
    from nanoj_esrrf_amd import _capi
    st = stack(4, w, h, seed=29)
    _, _, g_calc = run_engine(pkg, st, M, 1.5, 2)
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(w, h, M, 1.5, 2, 4, 4, 20000, e.allCLdevices[0], True, True, True, None, 0, 1)
    e.resetFramePosition()
    e.prepareDataFHT_SRRF(st)
    g = e.readGradientBuffers()
    assert e.frameCounters() == [0, 0]
    np.testing.assert_array_equal(g, g_calc)
    _, _, g_or = run_oracle(O, st, M, 1.5, 2)
    scale = float(np.abs(g_or).max())
    assert float(np.abs(g.reshape(-1) - np.asarray(g_or).reshape(-1)).max()) <= 1e-5 * scale
    # float frames take the same path
    e.prepareDataFHT_SRRF(st.astype(np.float32))
    np.testing.assert_array_equal(e.readGradientBuffers(), g_calc)
    e.release()
    # BIC mode has no FHT data path
    b = pkg.LiveSRRF_CL()
    b.checkDevices()
    b.initialise(w, h, 2, 1.5, 2, 4, 4, 20000, b.allCLdevices[0], True, True, False, None, 0, 1)
    with pytest.raises(_capi.EsrrfError):
        b.prepareDataFHT_SRRF(st)
    b.release()
