"""3-D (live3DSRRF.cl) parity of the HIP engine against the CPU oracle: montage split,
alignment, 3-D gradients, lateral interpolation, 3-D RGC, axial cubic, cumulants, 3-D MP.

Tolerance: per channel, both images divided by the oracle image's max |value|,
|a - b| <= 1e-5 + 1e-4 |b|.  Frame counters: exact.
"""
import numpy as np
import pytest

from conftest import normalised_close

pytestmark = pytest.mark.gpu


def planes_stack(n, w, h, nP, seed):
    from nanoj_esrrf_amd import synth
    return synth.blinking_stack_3d(n, w, h, nP, n_emitters=max(40, w * h // 5), n_filaments=4,
                                   seed=seed)


def montage(planes, rois, W=None, H=None):
    """Place plane z at montage cell rois[z] using LiveSRRF_CL's cell origins."""
    n, nP, h, w = planes.shape
    W = W or 3 * w
    H = H or 3 * h
    out = np.zeros((n, H, W), planes.dtype)
    for z in range(nP):
        i = int(rois[z]) % 3
        j = (int(rois[z]) - i) // 3
        x = int(np.floor(np.float32(W) / np.float32(3) * np.float32(i) + np.float32(0.5)))
        y = int(np.floor(np.float32(H) / np.float32(3) * np.float32(j) + np.float32(0.5)))
        out[:, y:y + h, x:x + w] = planes[:, z]
    return out


def calib_dict(sx, sy, th, roi, ic, bg):
    return {"X-shift (pixels)": sx, "Y-shift (pixels)": sy, "Theta (degrees)": th,
            "ROI #": roi, "Intensity scaling": ic, "Background level (ADC)": [bg]}


def run_both(pkg, O, mont, nP, M, fwhm, S, calib, intW=True, mp=True, drift=None, loads=None,
             nOnGPU=None, extra=None, Mz=0):
    n, H, W = mont.shape
    nOnGPU = nOnGPU or n
    cd = calib_dict(*calib)
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(W, H, M, fwhm, S, nOnGPU, n, 20000, e.allCLdevices[0], intW, mp, False, cd,
                 0, 1, axialMagnification=Mz)
    e.resetFramePosition()
    dx, dy = (np.zeros(n), np.zeros(n)) if drift is None else drift
    e.loadDriftXYGPUbuffer(dx, dy)
    o = O.OracleLiveSRRF(W, H, M, fwhm, S, nOnGPU, n, intW, mp, False,
                         calib3d=O.Calib3D(*[np.asarray(c, np.float64) for c in calib[:5]],
                                           float(calib[5])), axialMagnification=Mz or None)
    if drift is not None:
        o.loadDriftXYGPUbuffer(*drift)
    got_extra = {}
    for a, b in (loads or [(0, n)]):
        e.calculateSRRF(mont[a:b])
        o.calculateSRRF(mont[a:b].astype(np.float32))
        if extra and a == 0:
            got_extra["grad"] = (e.readGradientBuffers(), o.readGradientBuffers())
            got_extra["aligned"] = (e.readAlignedPixels(), o.readAlignedPixels())
    got = e.readSRRFbuffer()
    cnt = e.frameCounters()
    exp = o.readSRRFbuffer()
    if mp:
        got_extra["mp"] = (e.readMPmaps(), o.readMPmaps())
    e.release()
    return got, exp, cnt, list(o.counters), got_extra


def assert_parity(got, exp, what):
    got = np.asarray(got).reshape(4, -1)
    exp = np.asarray(exp).reshape(4, -1)
    for c, name in enumerate(("AVG", "VAR", "TAC2", "INT")):
        ok, ex, idx = normalised_close(got[c], exp[c])
        assert ok, f"{what}: {name} differs (excess {ex:.3g} at {idx})"


def zero_calib(nP):
    z = np.zeros(nP)
    return (z, z, z, np.arange(nP, dtype=np.float64), np.ones(nP), 0.0)


@pytest.mark.parametrize("w,h,nP,M,fwhm,S", [
    (16, 16, 4, 2, 1.5, 2), (12, 14, 3, 3, 1.5, 1), (20, 16, 9, 4, 1.5, 2),
    (16, 12, 5, 2, 2.0, 1.5),
])
def test_3d_zero_calibration_parity(pkg, O, w, h, nP, M, fwhm, S):
    planes = planes_stack(4, w, h, nP, seed=40 + w + nP)
    mont = montage(planes, np.arange(nP))
    got, exp, cnt, ocnt, _ = run_both(pkg, O, mont, nP, M, fwhm, S, zero_calib(nP))
    assert cnt == ocnt == [4, 4]
    assert_parity(got, exp, f"3D {w}x{h}x{nP} M{M} fwhm{fwhm} S{S}")


@pytest.mark.parametrize("w,h,nP,M,fwhm,loads", [
    (20, 16, 9, 4, 1.5, [(0, 11)]), (16, 12, 5, 2, 2.0, [(0, 5), (5, 13)]),
    (12, 14, 3, 3, 1.5, [(0, 1), (1, 9)]), (13, 9, 4, 4, 1.0, [(0, 3), (3, 4), (4, 7)]),
    (40, 36, 4, 3, 2.0, [(0, 6)]), (24, 24, 5, 5, 2.5, [(0, 2), (2, 5)]),
])
def test_3d_fast_kernels_match_window_kernel(pkg, monkeypatch, w, h, nP, M, fwhm, loads):
    """The drift-free 3-D kernels against the window-order kernel (k3d_rgc) on the same data:
    k3d_rgc2 (per-voxel geometry once per chunk of 8 frames), the tiled k3d_rgc3 and its
    two-z-slice form k3d_rgc3z (one staged tile for two slices, reciprocal multiplies)
    (wave-uniform tap tables per x/y class, z-slice and phase; LDS-staged gradient tiles;
    partial tiles, border classes, M = 2..5, one-frame and ragged loads continued across
    launches).  They sum the samples in different orders, so they agree to float rounding,
    not bit for bit: within the parity tolerance |a - b| <= 1e-5 + 1e-4 |b| on each
    channel divided by its maximum (TAC2 and VAR are products of per-frame values)."""
    n = loads[-1][1]
    planes = planes_stack(n, w, h, nP, seed=70 + w + nP)
    mont = montage(planes, np.arange(nP))
    cd = calib_dict(*zero_calib(nP))
    res = {}
    for kern, name in (("1", "k3d_rgc"), ("2", "k3d_rgc2"), ("3", "k3d_rgc3"),
                       ("3z", "k3d_rgc3z")):
        monkeypatch.setenv("ESRRF_3D_KERNEL", kern[0])
        monkeypatch.setenv("ESRRF_3D_ZPAIR", "1" if kern == "3z" else "0")
        e = pkg.LiveSRRF_CL()
        e.checkDevices()
        e.initialise(3 * w, 3 * h, M, fwhm, 2, max(b - a for a, b in loads), n, 20000,
                     e.allCLdevices[0], True, True, False, cd, 0, 1)
        e.resetFramePosition()
        e.loadDriftXYGPUbuffer(np.zeros(n), np.zeros(n))
        for a, b in loads:
            e.calculateSRRF(mont[a:b])
        assert e._e().rgc_kernel() == name
        res[kern] = np.asarray(e.readSRRFbuffer()).reshape(4, -1)
        e.release()
    for kern in ("2", "3", "3z"):
        for c in range(4):
            scale = max(float(np.abs(res["1"][c]).max()), 1e-30)
            b = res["1"][c] / scale
            excess = np.abs(res[kern][c] / scale - b) - (1e-5 + 1e-4 * np.abs(b))
            assert excess.max() <= 0, \
                f"kernel {kern} channel {c}: excess {excess.max():.3g} (scale {scale:.3g})"


def test_3d_calibrated_montage_parity(pkg, O):
    """Rotation, sub-pixel shifts, permuted ROI cells, intensity scaling, background, and a
    montage whose width and height are not multiples of 3."""
    w, h, nP, M = 16, 15, 6, 2
    planes = planes_stack(4, w, h, nP, seed=71)
    rng = np.random.default_rng(5)
    rois = rng.permutation(9)[:nP].astype(np.float64)
    mont = montage(planes, rois, W=3 * w + 1, H=3 * h + 2)
    calib = (rng.uniform(-0.7, 0.7, nP), rng.uniform(-0.7, 0.7, nP), rng.uniform(-4, 4, nP),
             rois, rng.uniform(0.8, 1.25, nP), 90.0)
    got, exp, _, _, ex = run_both(pkg, O, mont, nP, M, 1.5, 2, calib, extra=True)
    assert_parity(got, exp, "3D calibrated")
    ga, oa = ex["aligned"]
    np.testing.assert_allclose(ga.reshape(-1), oa.reshape(-1), rtol=1e-5,
                               atol=1e-5 * np.abs(oa).max())
    gg, og = ex["grad"]
    scale = np.abs(og).max()
    assert np.abs(gg.reshape(-1) - og.reshape(-1)).max() <= 1e-5 * scale
    gm, om = ex["mp"]
    np.testing.assert_allclose(gm.reshape(-1), om.reshape(-1), rtol=1e-4)


def test_3d_no_intensity_weighting_no_mp(pkg, O):
    w, h, nP, M = 14, 14, 4, 3
    planes = planes_stack(5, w, h, nP, seed=12)
    mont = montage(planes, np.arange(nP))
    got, exp, _, _, _ = run_both(pkg, O, mont, nP, M, 1.5, 1, zero_calib(nP), intW=False,
                                 mp=False)
    assert_parity(got, exp, "3D intW off, MP off")


def test_3d_uneven_loads_and_counters(pkg, O):
    w, h, nP, M = 12, 12, 3, 2
    planes = planes_stack(11, w, h, nP, seed=13)
    mont = montage(planes, np.arange(nP))
    loads = [(0, 4), (4, 8), (8, 11)]
    got, exp, cnt, ocnt, _ = run_both(pkg, O, mont, nP, M, 1.5, 2, zero_calib(nP), loads=loads,
                                      nOnGPU=4)
    assert cnt == ocnt == [11, 3]
    assert_parity(got, exp, "3D loads 4+4+3")


def test_3d_drift(pkg, O):
    w, h, nP, M = 14, 12, 4, 2
    planes = planes_stack(6, w, h, nP, seed=14)
    mont = montage(planes, np.arange(nP))
    rng = np.random.default_rng(8)
    drift = (rng.uniform(-0.3, 0.3, 6).astype(np.float32),
             rng.uniform(-0.3, 0.3, 6).astype(np.float32))
    got, exp, _, _, _ = run_both(pkg, O, mont, nP, M, 1.5, 2, zero_calib(nP), drift=drift)
    assert_parity(got, exp, "3D drift")


def test_3d_single_plane_is_rejected(pkg):
    from nanoj_esrrf_amd import _capi
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    with pytest.raises(_capi.EsrrfError):
        e.initialise(48, 48, 2, 1.5, 1, 2, 2, 20000, e.allCLdevices[0], True, True, False,
                     calib_dict(*zero_calib(1)), 0, 1)


@pytest.mark.parametrize("w,h,nP,M,Mz,loads", [
    (16, 16, 4, 2, 1, [(0, 4)]), (12, 14, 3, 3, 2, [(0, 2), (2, 5)]),
    (20, 16, 9, 4, 2, [(0, 4)]), (16, 12, 5, 2, 6, [(0, 3)]),
])
def test_3d_axial_magnification_parity(pkg, O, w, h, nP, M, Mz, loads):
    """Axial magnification Mz != M (no counterpart in the reference, whose output has nP * M
    planes): nP * Mz planes with centres (zM + 0.5) / Mz, M x M x Mz macro-voxels in the MP
    map.  The engine against the oracle's per-voxel restatement, MP maps included; one test
    each for a reduced (Mz < M), enlarged (Mz > M) and unit axial factor, with ragged
    loads."""
    n = loads[-1][1]
    planes = planes_stack(n, w, h, nP, seed=90 + w + nP + Mz)
    mont = montage(planes, np.arange(nP))
    got, exp, cnt, ocnt, ex = run_both(pkg, O, mont, nP, M, 1.5, 2, zero_calib(nP),
                                       loads=loads, nOnGPU=max(b - a for a, b in loads), Mz=Mz)
    assert np.asarray(got).shape == (4 * nP * Mz, h * M, w * M)
    assert cnt == ocnt == [n, loads[-1][1] - loads[-1][0]]
    assert_parity(got, exp, f"3D {w}x{h}x{nP} M{M} Mz{Mz}")
    gmp, omp = ex["mp"]
    assert gmp.shape == omp.shape == (3 * Mz, M, M)
    np.testing.assert_allclose(gmp, omp, rtol=1e-4, atol=1e-6)


def test_3d_axial_magnification_plane_invariant(pkg):
    """Plane 3z+1 at axial factor 3*Mz has the centre (3z + 1.5) / (3 Mz) = (z + 0.5) / Mz of
    plane z at Mz, so without MP correction the engine's AVG, VAR, TAC2 and INT planes agree
    exactly: every sample, weight and sum is the same float computation."""
    w, h, nP, M, n = 16, 12, 4, 2, 5
    planes = planes_stack(n, w, h, nP, seed=123)
    mont = montage(planes, np.arange(nP))
    cd = calib_dict(*zero_calib(nP))
    out = {}
    for Mz in (1, 3):
        e = pkg.LiveSRRF_CL()
        e.checkDevices()
        e.initialise(3 * w, 3 * h, M, 1.5, 2, n, n, 20000, e.allCLdevices[0], True, False,
                     False, cd, 0, 1, axialMagnification=Mz)
        e.resetFramePosition()
        e.loadDriftXYGPUbuffer(np.zeros(n), np.zeros(n))
        e.calculateSRRF(mont)
        out[Mz] = np.asarray(e.readSRRFbuffer()).reshape(4, nP * Mz, h * M, w * M)
        e.release()
    np.testing.assert_array_equal(out[3][:, 1::3], out[1])


def test_3d_reshape_matches_oracle_and_keeps_state(pkg, O):
    """reshapeImageStack3D (LiveSRRF_CL.java:963-994) on the device: permuted ROI cells, a
    montage whose sides are not multiples of 3, background and intensity scaling; uint16 and
    float frames.  Calling it between loads leaves the SR result unchanged."""
    w, h, nP, M = 14, 13, 5, 2
    planes = planes_stack(6, w, h, nP, seed=83)
    rng = np.random.default_rng(17)
    rois = rng.permutation(9)[:nP].astype(np.float64)
    mont = montage(planes, rois, W=3 * w + 2, H=3 * h + 1)
    n, H, W = mont.shape
    calib = (rng.uniform(-0.5, 0.5, nP), rng.uniform(-0.5, 0.5, nP), rng.uniform(-3, 3, nP),
             rois, rng.uniform(0.8, 1.25, nP), 75.0)
    o = O.OracleLiveSRRF(W, H, M, 1.5, 2, n, n, True, True, False,
                         calib3d=O.Calib3D(*[np.asarray(c, np.float64) for c in calib[:5]],
                                           float(calib[5])))
    exp = o.reshape3D(mont.astype(np.float32))
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(W, H, M, 1.5, 2, n, n, 20000, e.allCLdevices[0], True, True, False,
                 calib_dict(*calib), 0, 1)
    e.resetFramePosition()
    e.loadDriftXYGPUbuffer(np.zeros(n), np.zeros(n))
    got = e.reshapeImageStack3D(mont)
    assert got.shape == (n, nP, h, w)
    np.testing.assert_allclose(got, exp, rtol=1e-6, atol=0)
    np.testing.assert_allclose(e.reshapeImageStack3D(mont.astype(np.float32)), exp, rtol=1e-6,
                               atol=0)
    e.calculateSRRF(mont[:3])
    e.reshapeImageStack3D(mont[3:])
    e.calculateSRRF(mont[3:])
    res = e.readSRRFbuffer()
    e.release()
    ref, _, _, _, _ = run_both(pkg, O, mont, nP, M, 1.5, 2, calib, loads=[(0, 3), (3, n)])
    np.testing.assert_array_equal(res, ref)
