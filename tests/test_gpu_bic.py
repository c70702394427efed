"""BIC-mode (liveSRRF.cl) parity of the HIP engine against the CPU oracle.

Tolerance (the port's contract): per channel, both images divided by the oracle image's
max |value|, |a - b| <= 1e-5 + 1e-4 |b|.  Frame counters and schedules: exact.
"""
import numpy as np
import pytest

from conftest import normalised_close

pytestmark = pytest.mark.gpu


def stack(n, w, h, seed, emitters=None, fil=4):
    from nanoj_esrrf_amd import synth
    return synth.blinking_stack(n, w, h, emitters or max(20, w * h // 14), fil, seed=seed)


def run_engine(pkg, st, M, fwhm, S, nOnGPU=None, nFor=None, intW=True, mp=True, drift=None,
               loads=None, shape=None):
    """shape: launch-shape fields to pin (nw, taps_in_lds, frames_per_pass)."""
    n = st.shape[0]
    nOnGPU = nOnGPU or n
    nFor = nFor or n
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(st.shape[2], st.shape[1], M, fwhm, S, nOnGPU, nFor, 20000,
                 e.allCLdevices[0], intW, mp, False, None, 0, 1)
    if shape:
        e.setLaunchShape(**shape)
    e.resetFramePosition()
    dx, dy = (np.zeros(nFor), np.zeros(nFor)) if drift is None else drift
    e.loadDriftXYGPUbuffer(dx, dy)
    for a, b in (loads or [(0, n)]):
        e.calculateSRRF(st[a:b])
    out = e.readSRRFbuffer()
    counters = e.frameCounters()
    mpm = e.readMPmaps() if mp else None
    e.release()
    return out, counters, mpm


def run_oracle(O, st, M, fwhm, S, nOnGPU=None, nFor=None, intW=True, mp=True, drift=None,
               loads=None):
    n = st.shape[0]
    o = O.OracleLiveSRRF(st.shape[2], st.shape[1], M, fwhm, S, nOnGPU or n, nFor or n, intW, mp,
                         False)
    if drift is not None:
        o.loadDriftXYGPUbuffer(*drift)
    for a, b in (loads or [(0, n)]):
        o.calculateSRRF(st[a:b].astype(np.float32))
    return o.readSRRFbuffer(), list(o.counters), o.mp


def assert_parity(got, exp, what):
    for c, name in enumerate(("AVG", "VAR", "TAC2", "INT")):
        ok, ex, idx = normalised_close(got[c], exp[c])
        assert ok, f"{what}: {name} differs (excess {ex:.3g} at {idx})"


def test_plain_64x64x100_m5_parity(pkg, O):
    """The CPU-runnable workload: 100 frames of 64x64, M=5, fwhm 1.5, S=1, AVG + MP."""
    from nanoj_esrrf_amd import synth
    st = synth.plain_64x64x100()
    got, cnt, _ = run_engine(pkg, st, 5, 1.5, 1)
    exp, ocnt, _ = run_oracle(O, st, 5, 1.5, 1)
    assert cnt == ocnt == [100, 100]
    assert_parity(got.reshape(4, 320, 320), exp.reshape(4, 320, 320), "64x64x100 M5")


@pytest.mark.parametrize("M,fwhm,S,w,h", [
    (1, 1.5, 1, 21, 17), (2, 1.0, 2, 40, 24), (3, 2.0, 1, 19, 33), (4, 2.5, 2, 30, 30),
    (5, 3.0, 3, 27, 18), (5, 1.5, 2, 33, 40), (6, 1.2, 1, 16, 16), (8, 1.5, 2, 20, 14),
    (9, 1.5, 2, 12, 12), (5, 1.5, 1.7, 24, 24),
])
def test_bic_configs_parity(pkg, O, M, fwhm, S, w, h):
    st = stack(7, w, h, seed=100 + M * 7 + w)
    got, cnt, _ = run_engine(pkg, st, M, fwhm, S)
    exp, ocnt, _ = run_oracle(O, st, M, fwhm, S)
    assert cnt == ocnt
    assert_parity(got.reshape(4, h * M, w * M), exp.reshape(4, h * M, w * M),
                  f"M{M} fwhm{fwhm} S{S} {w}x{h}")


@pytest.mark.parametrize("M,w,h,loads", [
    (5, 40, 32, [(0, 5), (5, 9)]), (4, 27, 19, [(0, 7)]), (2, 24, 24, [(0, 1), (1, 4)]),
    (7, 16, 24, [(0, 6)]), (5, 33, 29, [(0, 2), (2, 13), (13, 16)]), (3, 20, 18, [(0, 11)]),
])
def test_bic_frame_pair_kernel_matches_single_frame(pkg, M, w, h, loads):
    """The frame-pair kernel (two frames per LDS pass) and the four-frame kernel (the default
    where the geometry fits it) reproduce the single-frame kernel bit for bit: tails of one
    to three frames, one-frame loads and continuation across loads included."""
    n = loads[-1][1]
    st = stack(n, w, h, seed=40 + M)
    res = {}
    for fp in (1, 2, 4):
        res[fp], _, _ = run_engine(pkg, st, M, 1.5, 2, nOnGPU=max(b - a for a, b in loads),
                                   nFor=n, loads=loads, shape=dict(frames_per_pass=fp))
    assert np.array_equal(res[1], res[2])
    assert np.array_equal(res[1], res[4])


def test_bic_four_frame_kernel_is_the_default(pkg):
    """Where the geometry fits it (parity-plane pitch 24, tile rows, LDS), the engine runs
    four frames per pass by default; a wider window falls back to frame pairs."""
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(40, 32, 5, 1.5, 2, 4, 4, 20000, e.allCLdevices[0], True, True, False, None, 0, 1)
    assert e.launchShape()["frames_per_pass"] == 4
    e.release()
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(64, 64, 5, 4.0, 2, 4, 4, 20000, e.allCLdevices[0], True, True, False, None, 0, 1)
    assert e.launchShape()["frames_per_pass"] in (1, 2)
    e.release()


@pytest.mark.parametrize("M,w,h,loads", [
    (5, 40, 32, [(0, 5), (5, 9)]), (4, 27, 19, [(0, 7)]), (3, 64, 48, [(0, 4)]),
])
def test_bic_frame_pair_taps_from_scalar_loads(pkg, M, w, h, loads):
    """The frame-pair kernel reading its tap tables with scalar loads from global memory
    reproduces the variant that stages them in LDS bit for bit (same operations, the tap
    values only come from SGPRs instead of VGPRs)."""
    n = loads[-1][1]
    st = stack(n, w, h, seed=70 + M)
    res = {}
    for v in (0, 1):
        res[v], _, _ = run_engine(pkg, st, M, 1.5, 2, nOnGPU=max(b - a for a, b in loads),
                                  nFor=n, loads=loads,
                                  shape=dict(frames_per_pass=2, taps_in_lds=v))
    assert np.array_equal(res[0], res[1])


@pytest.mark.parametrize("M,fwhm,w,h,nw,tl,fp", [(5, 1.5, 40, 32, "8", 1, 2),
                                                  (4, 2.0, 27, 19, "4", 1, 2),
                                                  (2, 1.0, 24, 24, "4", 1, 2),
                                                  (5, 3.0, 16, 16, "12", 0, 2),
                                                  (5, 1.5, 40, 32, "8", 1, 4),
                                                  (5, 1.5, 40, 32, "16", 1, 4),
                                                  (4, 2.0, 27, 19, "4", 1, 4)])
def test_bic_frame_pair_register_staging(pkg, monkeypatch, M, fwhm, w, h, nw, tl, fp):
    """Staging the next frame pair's (or four frames') tiles in registers during the current
    pass (stored to LDS between the barriers) gives the same bits as filling them after it,
    odd frame tails and continuation across loads included."""
    st = stack(7, w, h, seed=140 + M)
    res = {}
    for sg in ("0", "1"):
        monkeypatch.setenv("ESRRF_BIC_PAIR_DIRECT", sg)
        res[sg], _, _ = run_engine(pkg, st, M, fwhm, 2, nOnGPU=7, loads=[(0, 5), (5, 7)],
                                   shape=dict(frames_per_pass=fp, nw=int(nw), taps_in_lds=tl))
    assert np.array_equal(res["0"], res["1"])


@pytest.mark.parametrize("nw", ["7", "9", "13"])
def test_bic_frame_pair_balanced_wave_counts(pkg, nw):
    """Blocks of 7, 9 or 13 waves (the balanced shapes for 25 phases) give the same bits as
    the default 8-wave block: each (pixel, phase) is evaluated the same way by whichever
    wave holds it."""
    st = stack(6, 40, 32, seed=91)
    res = {}
    for v in ("8", nw):
        res[v], _, _ = run_engine(pkg, st, 5, 1.5, 2, nOnGPU=6, loads=[(0, 6)],
                                  shape=dict(nw=int(v)))
    assert np.array_equal(res["8"], res[nw])


@pytest.mark.parametrize("M,w,h", [(3, 4, 4), (2, 5, 4), (5, 4, 9)])
def test_bic_minimum_frame_sizes(pkg, O, M, w, h):
    """Frames at the 4x4 minimum: every output pixel is a border pixel (bilinear
    extrapolation, gated windows cut by both edges) and one tile overhangs the image."""
    st = stack(5, w, h, seed=17 + w * h, emitters=6, fil=1)
    got, cnt, _ = run_engine(pkg, st, M, 1.5, 2)
    exp, ocnt, _ = run_oracle(O, st, M, 1.5, 2)
    assert cnt == ocnt
    assert_parity(got.reshape(4, h * M, w * M), exp.reshape(4, h * M, w * M),
                  f"M{M} {w}x{h}")


@pytest.mark.parametrize("w,h,M", [(3, 8, 2), (8, 3, 2), (8, 8, 0), (0, 8, 2)])
def test_bic_rejects_invalid_arguments(pkg, w, h, M):
    """Frames under 4x4 pixels and a magnification below 1 are refused at initialise
    with the engine's error, before any device memory is allocated."""
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    with pytest.raises(pkg._capi.EsrrfError):
        e.initialise(w, h, M, 1.5, 2, 4, 4, 20000, e.allCLdevices[0], True, True, False, None,
                     0, 1)


def test_bic_cancel_between_chunks(pkg):
    """A cancel flag raised before calculateSRRF (the plugin's Esc key, LiveSRRF_CL.java:731)
    stops the load after its first chunk of 16 frames: calculateSRRF reports the
    cancellation, the counters stand at 16, and the state equals a load of exactly those 16
    frames."""
    st = stack(40, 24, 20, seed=33)
    res = {}
    for mode in ("cancel", "sixteen"):
        e = pkg.LiveSRRF_CL()
        e.checkDevices()
        e.initialise(24, 20, 3, 1.5, 2, 40, 40, 20000, e.allCLdevices[0], True, True, False,
                     None, 0, 1)
        e.resetFramePosition()
        e.loadDriftXYGPUbuffer(np.zeros(40), np.zeros(40))
        if mode == "cancel":
            assert e.calculateSRRF(st, cancel=lambda: True) is True
        else:
            assert e.calculateSRRF(st[:16]) is False
        assert list(e.frameCounters()) == [16, 16]
        res[mode] = np.asarray(e.readSRRFbuffer()).copy()
        e.release()
    assert np.array_equal(res["cancel"], res["sixteen"])


def test_bic_single_frame_sr_image(pkg, O):
    """nFrameForSRRF = 1: TAC2's 1/(n - 1) is a division by zero in the reference
    (liveSRRF.cl:374,377); the first frame initialises TAC2 to 0, so the result is finite:
    AVG, TAC2 (= -AVG^2) and INT equal the oracle's, and VAR = E2 - AVG^2 is rounding
    noise below 1e-5 of AVG^2 in both (without the MP correction, which would divide that
    noise by its own macro-pixel mean)."""
    st = stack(1, 20, 16, seed=21)
    got, cnt, _ = run_engine(pkg, st, 4, 1.5, 2, mp=False)
    exp, ocnt, _ = run_oracle(O, st, 4, 1.5, 2, mp=False)
    assert cnt == ocnt == [1, 1]
    assert np.isfinite(got).all()
    g, x = got.reshape(4, 64, 80), exp.reshape(4, 64, 80)
    for c, name in ((0, "AVG"), (2, "TAC2"), (3, "INT")):
        ok, ex, idx = normalised_close(g[c], x[c])
        assert ok, f"nF=1: {name} differs (excess {ex:.3g} at {idx})"
    bound = 1e-5 * float(np.max(x[0].astype(np.float64) ** 2))
    assert np.abs(g[1]).max() <= bound and np.abs(x[1]).max() <= bound


def test_bic_allocation_failure_is_reported(pkg):
    """A frame budget the device cannot hold (1e7 frames of 256 x 256) is refused by
    initialise with the engine's error, and the device stays usable."""
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    with pytest.raises(pkg._capi.EsrrfError):
        e.initialise(256, 256, 5, 1.5, 2, 10_000_000, 10_000_000, 20000, e.allCLdevices[0],
                     True, True, False, None, 0, 1)
    st = stack(3, 16, 16, seed=2)
    got, cnt, _ = run_engine(pkg, st, 2, 1.5, 2)
    assert cnt == [3, 3] and np.isfinite(got).all()


def test_bic_no_intensity_weighting(pkg, O):
    st = stack(6, 24, 20, seed=5)
    got, _, _ = run_engine(pkg, st, 4, 1.5, 2, intW=False)
    exp, _, _ = run_oracle(O, st, 4, 1.5, 2, intW=False)
    assert_parity(got.reshape(4, 80, 96), exp.reshape(4, 80, 96), "intWeighting off")


def test_bic_uneven_loads_and_counters(pkg, O):
    """One SR frame of 11 frames pushed as loads of 4, 4, 3 (gui/LiveSRRF_optimised_.java
    :309-321): accumulators persist across launches, counters stay exact."""
    st = stack(11, 24, 24, seed=11)
    loads = [(0, 4), (4, 8), (8, 11)]
    got, cnt, _ = run_engine(pkg, st, 5, 1.5, 2, nOnGPU=4, nFor=11, loads=loads)
    exp, ocnt, _ = run_oracle(O, st, 5, 1.5, 2, nOnGPU=4, nFor=11, loads=loads)
    assert cnt == ocnt == [11, 3]
    assert_parity(got.reshape(4, 120, 120), exp.reshape(4, 120, 120), "uneven loads")


def test_bic_drift_generic_path(pkg, O):
    """Per-frame drift (x M on upload, LiveSRRF_CL.java:587-588) takes the per-pixel path."""
    st = stack(5, 20, 20, seed=3)
    rng = np.random.default_rng(0)
    dx = rng.uniform(-0.3, 0.3, 5).astype(np.float32)
    dy = rng.uniform(-0.3, 0.3, 5).astype(np.float32)
    got, _, _ = run_engine(pkg, st, 3, 1.5, 1, drift=(dx, dy))
    exp, _, _ = run_oracle(O, st, 3, 1.5, 1, drift=(dx, dy))
    assert_parity(got.reshape(4, 60, 60), exp.reshape(4, 60, 60), "drift")


def test_bic_mp_map_and_gradients(pkg, O):
    st = stack(4, 18, 22, seed=9)
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(18, 22, 3, 1.5, 1, 4, 4, 20000, e.allCLdevices[0], True, True, False, None, 0, 1)
    e.resetFramePosition()
    e.calculateSRRF(st)
    g = e.readGradientBuffers()
    e.readSRRFbuffer()
    mp = e.readMPmaps()
    e.release()
    o = O.OracleLiveSRRF(18, 22, 3, 1.5, 1, 4, 4, True, True, False)
    o.calculateSRRF(st.astype(np.float32))
    og = o.readGradientBuffers()
    o.readSRRFbuffer()
    # gradients of integer data are exact (Catmull-Rom half-point weights are dyadic)
    assert np.array_equal(g, og)
    np.testing.assert_allclose(mp.reshape(-1), o.mp, rtol=1e-4)


def test_bic_constant_stack(pkg):
    """Constant frames: zero gradient -> CGLH = 0 -> AVG = VAR = TAC2 = 0, INT = value."""
    st = np.full((5, 16, 16), 123, np.uint16)
    got, _, _ = run_engine(pkg, st, 3, 1.5, 1, mp=False)
    got = got.reshape(4, 48, 48)
    assert np.all(got[0] == 0) and np.all(got[1] == 0) and np.all(got[2] == 0)
    np.testing.assert_allclose(got[3], 123.0, rtol=1e-6)


@pytest.mark.parametrize("w,h,M,loads", [
    (64, 64, 5, [(0, 6)]), (45, 37, 3, [(0, 2), (2, 5)]), (7, 9, 2, [(0, 4)]),
    (100, 33, 4, [(0, 3)]), (40, 36, 5, [(0, 1), (1, 8), (8, 11)]),
])
def test_bic_prep_shared_row_sums_match_per_point_stencils(pkg, monkeypatch, w, h, M, loads):
    """k_bic_prep4 (four frames per block, the default), k_bic_prep2 (row sums shared per
    block) and k_bic_prep (a 4x4 stencil per sample point) feed the RGC the same unit
    gradients: the results agree bit for bit (tiles that overhang the image, border points on
    the bilinear-extrapolation branch, tiny frames, loads of 1 to 3 frames past a multiple
    of four)."""
    n = loads[-1][1]
    st = stack(n, w, h, seed=300 + w)
    res = {}
    for v in ("1", "2", "4"):
        monkeypatch.setenv("ESRRF_BIC_PREP", v)
        res[v], _, _ = run_engine(pkg, st, M, 1.5, 2, nOnGPU=max(b - a for a, b in loads),
                                  loads=loads)
    assert np.array_equal(res["1"], res["2"])
    assert np.array_equal(res["1"], res["4"])


@pytest.mark.parametrize("n,chunks", [(130, 1), (130, 2), (301, 4), (301, 16), (257, 3)])
def test_pipelined_host_upload_bit_identical(pkg, n, chunks):
    """esrrf_calculate splits a load of host frames into chunks whose upload overlaps the
    previous chunk's kernels; each chunk's RGC launch continues the cumulants of the last.
    The result, counters included, equals the one-copy path and the device-resident path
    bit for bit, for pageable and pinned host memory and odd chunk boundaries."""
    import ctypes
    from conftest import hip_runtime
    hip = hip_runtime()
    C = pkg._capi
    w, h, M = 40, 36, 5
    st = stack(n, w, h, seed=900 + n).astype(np.uint16)
    p = C.Params()
    p.width, p.height, p.magnification, p.fwhm, p.sensitivity = w, h, M, 1.5, 2.0
    p.n_frames_on_gpu, p.n_frame_for_srrf, p.block_length, p.device = n, n, 20000, 0
    p.int_weighting, p.do_mp_map_correction, p.do_fht_interpolation = 1, 1, 0
    buf = ctypes.c_void_p()

    def run(mode):
        eng = pkg.Engine(p)
        L, hd = eng.L, eng.h
        out = np.zeros(4 * w * M * h * M, np.float32)
        eng.check(L.esrrf_set_upload_chunks(hd, 1 if mode == "single" else chunks))
        eng.check(L.esrrf_reset_frame_position(hd))
        if mode == "device":
            assert hip.hipMalloc(ctypes.byref(buf), st.nbytes) == 0
            assert hip.hipMemcpy(buf, st.ctypes.data, st.nbytes, 1) == 0   # host to device
            eng.check(L.esrrf_calculate_device(hd, buf, C.ESRRF_U16, n, None, None))
        elif mode == "pinned":
            hb = C.HostBuffer(st.nbytes)
            hb.array(np.uint16, st.shape)[...] = st
            eng.check(L.esrrf_calculate(hd, hb.ptr, C.ESRRF_U16, n, None, None))
        else:
            eng.check(L.esrrf_calculate(hd, st.ctypes.data_as(ctypes.c_void_p), C.ESRRF_U16,
                                        n, None, None))
        eng.check(L.esrrf_read_result(hd, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        cnt = eng.counters()
        eng.close()
        if mode == "device":
            hip.hipFree(buf)
        elif mode == "pinned":
            hb.close()
        return out, cnt

    ref, rcnt = run("single")
    assert rcnt == [n, n]
    for mode in ("pageable", "pinned", "device"):
        got, cnt = run(mode)
        assert cnt == rcnt, mode
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), mode


@pytest.mark.parametrize("fht,mp,const", [(0, 1, False), (0, 0, False), (1, 0, False),
                                          (0, 1, True)])
def test_host_read_back_matches_device_read_back(pkg, fht, mp, const):
    """esrrf_read_result copies the INT channel to the host on a second stream while VAR and
    the MP correction finish the other three; the result equals esrrf_read_result_device's
    (one NaN-to-0 pass over all four channels after the epilogue) bit for bit, with and
    without the MP correction, in FHT mode, and for a constant stack whose CGLH is 0/0."""
    import ctypes
    from conftest import hip_runtime
    hip = hip_runtime()
    C = pkg._capi
    n, w, h, M = 24, 40, 36, 4
    st = stack(n, w, h, seed=950 + 2 * fht + mp).astype(np.uint16)
    if const:
        st[...] = 300
    p = C.Params()
    p.width, p.height, p.magnification, p.fwhm, p.sensitivity = w, h, M, 1.5, 2.0
    p.n_frames_on_gpu, p.n_frame_for_srrf, p.block_length, p.device = n, n, 20000, 0
    p.int_weighting, p.do_mp_map_correction, p.do_fht_interpolation = 1, mp, fht
    nv = 4 * w * M * h * M

    def run(device):
        eng = pkg.Engine(p)
        L, hd = eng.L, eng.h
        eng.check(L.esrrf_reset_frame_position(hd))
        eng.check(L.esrrf_calculate(hd, st.ctypes.data_as(ctypes.c_void_p), C.ESRRF_U16, n,
                                    None, None))
        out = np.zeros(nv, np.float32)
        if device:
            buf = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(buf), out.nbytes) == 0
            eng.check(L.esrrf_read_result_device(hd, buf))
            eng.check(L.esrrf_finish(hd))
            assert hip.hipMemcpy(out.ctypes.data, buf, out.nbytes, 2) == 0   # device to host
            hip.hipFree(buf)
        else:
            eng.check(L.esrrf_read_result(hd, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        eng.close()
        return out

    host, dev = run(False), run(True)
    assert not np.isnan(host).any()
    assert np.array_equal(host.view(np.uint32), dev.view(np.uint32))
    if const:
        assert np.all(host[: 3 * nv // 4] == 0.0)
        assert np.allclose(host[3 * nv // 4:], 300.0, rtol=1e-5, atol=0)


def test_upload_chunks_refuses_out_of_range(pkg):
    C = pkg._capi
    p = C.Params()
    p.width, p.height, p.magnification, p.fwhm, p.sensitivity = 16, 16, 2, 1.5, 1.0
    p.n_frames_on_gpu, p.n_frame_for_srrf, p.block_length, p.device = 4, 4, 20000, 0
    p.int_weighting, p.do_mp_map_correction, p.do_fht_interpolation = 1, 1, 0
    eng = pkg.Engine(p)
    ESRRF_ERR_INVALID = 1
    assert eng.L.esrrf_set_upload_chunks(eng.h, 0) == ESRRF_ERR_INVALID
    assert eng.L.esrrf_set_upload_chunks(eng.h, 17) == ESRRF_ERR_INVALID
    assert eng.L.esrrf_set_upload_chunks(eng.h, 16) == C.ESRRF_OK
    eng.close()
