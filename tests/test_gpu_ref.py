"""The reference's own OpenCL kernels (resources/*.cl compiled from their sources into
oracle/_ref/ and run through the AMD OpenCL runtime with LiveSRRF_CL's enqueue order:
gradient -> interpolation -> per-frame RGC in 20000-item blocks -> increment; VAR -> MP)
pin the CPU oracle, and the HIP engine is checked against them directly.

Tolerance: per channel, normalised by the reference image's max |value|,
|a - b| <= 1e-5 + 1e-4 |b|.  Integer-valued gradients: exact.
"""
import zlib

import numpy as np
import pytest

from conftest import normalised_close

pytestmark = pytest.mark.gpu

refrun = pytest.importorskip("oracle.refrun")
if not refrun.available():
    pytest.skip("oracle/_ref not built (needs the reference sources at build time)",
                allow_module_level=True)


def stack(cfg, seed):
    from nanoj_esrrf_amd import synth
    mode, w, h, nP, M, fwhm, S, nF, iw = cfg
    if mode == "3d":
        return synth.blinking_stack_3d(nF, w, h, nP, n_emitters=60, n_filaments=4, seed=seed)
    return synth.blinking_stack(nF, w, h, max(20, w * h // 14), 4, seed=seed)


def assert_close(got, exp, what, nch=4):
    got = np.asarray(got).reshape(nch, -1)
    exp = np.asarray(exp).reshape(nch, -1)
    for c, name in enumerate(("AVG", "VAR", "TAC2", "INT")[:nch]):
        ok, ex, idx = normalised_close(got[c], exp[c])
        assert ok, f"{what}: {name} differs (excess {ex:.3g} at {idx})"


BIC = [k for k, v in refrun.configs().items() if v[0] == "bic"]
FHT = [k for k, v in refrun.configs().items() if v[0] == "fht"]
D3 = [k for k, v in refrun.configs().items() if v[0] == "3d"]


@pytest.mark.parametrize("name", BIC)
def test_bic_oracle_and_engine_match_reference_kernels(pkg, O, name):
    cfg = refrun.configs()[name]
    _, w, h, _, M, fwhm, S, nF, iw = cfg
    # a fixed seed per configuration (str hashes change from process to process)
    st = stack(cfg, seed=zlib.crc32(name.encode()) % 1000)
    loads = [4, 4, 3] if name.endswith("loads") else [nF]
    ref, ref_mp, gint = refrun.run(name, st.astype(np.float32), loads=loads, want_gint=True)
    # oracle
    o = O.OracleLiveSRRF(w, h, M, fwhm, S, max(loads), nF, bool(iw), True, False)
    a = 0
    for n in loads:
        o.calculateSRRF(st[a:a + n].astype(np.float32))
        if a == 0:
            og = o.readGradientBuffers()
        a += n
    exp = o.readSRRFbuffer()
    assert_close(exp, ref, f"oracle vs reference {name}")
    np.testing.assert_allclose(o.mp, ref_mp, rtol=1e-4)
    assert np.array_equal(og.reshape(-1), gint)      # integer data: exact
    # HIP engine
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(w, h, M, fwhm, S, max(loads), nF, 20000, e.allCLdevices[0], bool(iw), True,
                 False, None, 0, 1)
    e.resetFramePosition()
    e.loadDriftXYGPUbuffer(np.zeros(nF), np.zeros(nF))
    a = 0
    for n in loads:
        e.calculateSRRF(st[a:a + n])
        a += n
    got = e.readSRRFbuffer()
    mp = e.readMPmaps()
    e.release()
    assert_close(got, ref, f"HIP vs reference {name}")
    np.testing.assert_allclose(mp.reshape(-1), ref_mp, rtol=1e-4)


@pytest.mark.parametrize("name", BIC[:1])
def test_bic_drift_matches_reference_kernels(pkg, O, name):
    cfg = refrun.configs()[name]
    _, w, h, _, M, fwhm, S, nF, iw = cfg
    st = stack(cfg, seed=77)
    rng = np.random.default_rng(1)
    dx = rng.uniform(-0.2, 0.2, nF).astype(np.float32)
    dy = rng.uniform(-0.2, 0.2, nF).astype(np.float32)
    drift = np.concatenate([np.float32(M) * dx, np.float32(M) * dy]).astype(np.float32)
    ref, _, _ = refrun.run(name, st.astype(np.float32), drift=drift)
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(w, h, M, fwhm, S, nF, nF, 20000, e.allCLdevices[0], bool(iw), True, False,
                 None, 0, 1)
    e.resetFramePosition()
    e.loadDriftXYGPUbuffer(dx, dy)
    e.calculateSRRF(st)
    got = e.readSRRFbuffer()
    e.release()
    assert_close(got, ref, f"HIP vs reference with drift {name}")


@pytest.mark.parametrize("name", FHT)
def test_fht_oracle_and_engine_match_reference_kernels(pkg, O, name):
    from oracle import fht
    cfg = refrun.configs()[name]
    _, w, h, _, M, fwhm, S, nF, iw = cfg
    st = stack(cfg, seed=5)
    up = fht.fht_stack(st.astype(np.float32), M, True)
    ref, _, _ = refrun.run(name, up)
    o = O.OracleLiveSRRF(w, h, M, fwhm, S, nF, nF, bool(iw), True, True)
    o.calculateSRRF(st.astype(np.float32))
    exp = o.readSRRFbuffer()
    assert_close(exp, ref, f"oracle vs reference {name}")
    # HIP engine (its own FFT upsampling) against the reference kernels fed with the
    # oracle's upsampled frames
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(w, h, M, fwhm, S, nF, nF, 20000, e.allCLdevices[0], bool(iw), True, True,
                 None, 0, 1)
    e.resetFramePosition()
    e.loadDriftXYGPUbuffer(np.zeros(nF), np.zeros(nF))
    e.calculateSRRF(st)
    got = e.readSRRFbuffer()
    e.release()
    assert_close(got, ref, f"HIP vs reference {name}")


@pytest.mark.parametrize("name", D3)
def test_3d_oracle_matches_reference_kernels(O, name):
    cfg = refrun.configs()[name]
    _, w, h, nP, M, fwhm, S, nF, iw = cfg
    planes = stack(cfg, seed=9).astype(np.float32)          # [n][nP][h][w]
    from nanoj_esrrf_amd import synth
    mont = synth.montage3x3(planes)
    z = np.zeros(nP)
    calib = O.Calib3D(z, z, z, np.arange(nP, dtype=np.float64), np.ones(nP), 0.0)
    o = O.OracleLiveSRRF(3 * w, 3 * h, M, fwhm, S, nF, nF, bool(iw), True, False, calib3d=calib)
    o.calculateSRRF(mont)
    exp = o.readSRRFbuffer()
    ref, ref_mp, _ = refrun.run(name, planes, st3d=refrun.st3d_of(z, z, z))
    assert_close(exp, ref, f"oracle vs reference {name}")
    np.testing.assert_allclose(o.mp, ref_mp, rtol=1e-4)


@pytest.mark.parametrize("calibrated", [False, True])
@pytest.mark.parametrize("name", D3)
def test_3d_engine_matches_reference_kernels(pkg, name, calibrated):
    cfg = refrun.configs()[name]
    _, w, h, nP, M, fwhm, S, nF, iw = cfg
    planes = stack(cfg, seed=19).astype(np.float32)
    from nanoj_esrrf_amd import synth
    mont = synth.montage3x3(planes)
    rng = np.random.default_rng(4)
    if calibrated:
        sx, sy, th = rng.uniform(-0.6, 0.6, nP), rng.uniform(-0.6, 0.6, nP), rng.uniform(-3, 3, nP)
    else:
        sx = sy = th = np.zeros(nP)
    ref, ref_mp, _ = refrun.run(name, planes, st3d=refrun.st3d_of(th, sx, sy))
    calib = {"X-shift (pixels)": sx, "Y-shift (pixels)": sy, "Theta (degrees)": th,
             "ROI #": np.arange(nP, dtype=np.float64), "Intensity scaling": np.ones(nP),
             "Background level (ADC)": [0.0]}
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(3 * w, 3 * h, M, fwhm, S, nF, nF, 20000, e.allCLdevices[0], bool(iw), True,
                 False, calib, 0, 1)
    e.resetFramePosition()
    e.loadDriftXYGPUbuffer(np.zeros(nF), np.zeros(nF))
    e.calculateSRRF(mont)
    got = e.readSRRFbuffer()
    mp = e.readMPmaps()
    e.release()
    assert_close(got, ref, f"HIP vs reference {name} calibrated={calibrated}")
    np.testing.assert_allclose(mp.reshape(-1), ref_mp, rtol=1e-4)
