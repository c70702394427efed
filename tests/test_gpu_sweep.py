"""Fused parameter sweep (gui/ParametersSweep_.java:328-519) against the per-configuration
path it replaces: for every (magnification, frame count, sensitivity, fwhm) the sweep must
return what initialise + calculateSRRF(first nf frames) + readSRRFbuffer returns
(intensity weighting and MP correction on, as the reference sweep sets them, :452).

Tolerance: per channel, normalised by the reference image's max |value|,
|a - b| <= 1e-5 + 1e-4 |b| (the same contract as the plain engine's parity tests).
"""
import numpy as np
import pytest

from conftest import normalised_close

pytestmark = pytest.mark.gpu


def stack(n, w, h, seed):
    from nanoj_esrrf_amd import synth
    return synth.blinking_stack(n, w, h, max(20, w * h // 14), 4, seed=seed)


def plain(pkg, st, M, fwhm, S, nf):
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(st.shape[2], st.shape[1], M, fwhm, S, nf, nf, 20000, e.allCLdevices[0], True,
                 True, False, None, 0, 1)
    e.resetFramePosition()
    e.loadDriftXYGPUbuffer(np.zeros(nf), np.zeros(nf))
    e.calculateSRRF(st[:nf])
    out = e.readSRRFbuffer()
    e.release()
    return out


def assert_parity(got, exp, what):
    for c, name in enumerate(("AVG", "VAR", "TAC2", "INT")):
        ok, ex, idx = normalised_close(got[c], exp[c])
        assert ok, f"{what}: {name} differs (excess {ex:.3g} at {idx})"


MAGS, FWHM, SENS, NFS = [2, 3], [1.0, 1.5], [1.0, 2.0, 2.5], [5, 12]


@pytest.fixture(scope="module")
def swept(pkg):
    st = stack(14, 24, 20, seed=61)
    sw = pkg.ParametersSweep(24, 20, MAGS, FWHM, SENS, NFS, use_graph=False)
    sw.run(st)
    res = {(im, iq, i_s, i_f): sw.read(im, iq, i_s, i_f)
           for im in range(len(MAGS)) for iq in range(len(NFS))
           for i_s in range(len(SENS)) for i_f in range(len(FWHM))}
    sw.close()
    return st, res


def test_sweep_matches_per_configuration_engine(pkg, swept):
    st, res = swept
    for (im, iq, i_s, i_f), got in res.items():
        exp = plain(pkg, st, MAGS[im], FWHM[i_f], SENS[i_s], NFS[iq])
        assert_parity(got, exp, f"M{MAGS[im]} nf{NFS[iq]} S{SENS[i_s]} fwhm{FWHM[i_f]}")


@pytest.mark.parametrize("nfs", [[3, 5, 8, 12], [2, 4, 6, 9, 12]])
def test_sweep_many_frame_counts(pkg, nfs):
    """Four frame counts (two sensitivities per thread) and five (one per thread, every
    frame-count slot) accumulate what the per-configuration engine computes."""
    st = stack(12, 20, 16, seed=66)
    sens = [1.0, 3.0, 1.5]
    sw = pkg.ParametersSweep(20, 16, [3], [1.5], sens, nfs, use_graph=False)
    sw.run(st)
    for iq, nf in enumerate(nfs):
        for i_s, S in enumerate(sens):
            exp = plain(pkg, st, 3, 1.5, S, nf)
            assert_parity(sw.read(0, iq, i_s, 0), exp, f"nf{nf} S{S}")
    sw.close()


def test_sweep_matches_oracle(pkg, O, swept):
    st, res = swept
    for im, iq, i_s, i_f in [(0, 1, 1, 1), (1, 0, 2, 0)]:
        M, nf = MAGS[im], NFS[iq]
        o = O.OracleLiveSRRF(24, 20, M, FWHM[i_f], SENS[i_s], nf, nf, True, True, False)
        o.calculateSRRF(st[:nf].astype(np.float32))
        exp = o.readSRRFbuffer()
        assert_parity(res[(im, iq, i_s, i_f)], exp, f"oracle M{M} nf{nf}")


def test_sweep_graph_replays_match_eager(pkg):
    """The captured hipGraph is replayed several times, with eager work and host
    synchronisation between the replays, and on a second stack; every replay must equal
    the eager sequence bit for bit."""
    a, b = stack(9, 20, 16, seed=1), stack(9, 20, 16, seed=2)
    args = (20, 16, [2, 4], [1.5], [1.0, 2.0], [4, 9])
    eager = pkg.ParametersSweep(*args, use_graph=False)
    graph = pkg.ParametersSweep(*args, use_graph=True)
    for st in (a, b, a):
        eager.run(st)
        for rep in range(3):
            graph.run(st)
            for cfg in [(0, 0, 0, 0), (1, 1, 1, 0), (0, 1, 1, 0)]:
                assert np.array_equal(graph.read(*cfg), eager.read(*cfg)), (rep, cfg)
            plain(pkg, st, 2, 1.5, 1.0, 4)          # unrelated eager work between replays
    eager.close()
    graph.close()


def test_sweep_autotune_picks_a_supported_shape(pkg):
    st = stack(8, 24, 24, seed=5)
    tuned = pkg.ParametersSweep(24, 24, [5], [1.5], [2.0], [8], autotune=True, use_graph=False)
    tuned.run(st)
    sh = tuned.shape(0, 0)
    # emit blocks hold a group of nw x ppw of the 25 phases (grid.y covers the groups)
    assert sh["ms"] > 0 and sh["nw"] in (4, 8, 12, 16) and sh["ppw"] in (1, 2)
    assert sh["frames_per_pass"] in (1, 2)
    base = pkg.ParametersSweep(24, 24, [5], [1.5], [2.0], [8], autotune=False, use_graph=False)
    base.run(st)
    assert_parity(tuned.read(0, 0, 0, 0), base.read(0, 0, 0, 0), "autotuned vs default")
    tuned.close()
    base.close()


def test_sweep_rejects_short_stacks_and_large_magnification(pkg):
    from nanoj_esrrf_amd import _capi
    sw = pkg.ParametersSweep(16, 16, [2], [1.5], [1.0], [10], use_graph=False)
    with pytest.raises(_capi.EsrrfError):
        sw.run(stack(6, 16, 16, seed=3))
    sw.close()
    with pytest.raises(_capi.EsrrfError):
        pkg.ParametersSweep(16, 16, [9], [1.5], [1.0], [4])


def test_engine_autotune_keeps_results(pkg):
    st = stack(10, 32, 24, seed=9)
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(32, 24, 5, 1.5, 2, 10, 10, 20000, e.allCLdevices[0], True, True, False, None,
                 0, 1)
    e.resetFramePosition()
    e.loadDriftXYGPUbuffer(np.zeros(10), np.zeros(10))
    e.calculateSRRF(st)
    before = e.readSRRFbuffer()
    shape = e.autotune(reps=1)
    assert shape["ms"] > 0
    e.resetFramePosition()
    e.calculateSRRF(st)
    after = e.readSRRFbuffer()
    e.release()
    assert_parity(after, before, "after autotune")


# --- the batched-sweep workload's full parameter grid on a reduced stack -------------------
C3_MAGS, C3_FWHM, C3_SENS, C3_NFS = [2, 4, 5], [1.0, 1.5, 2.0, 2.5, 3.0], [1.0, 2.0, 3.0, 4.0], \
    [6, 15, 24]


def test_sweep_full_grid_autotuned_matches_per_configuration_engine(pkg):
    """The 180 configurations of magnification {2,4,5} x fwhm {1..3} x sensitivity {1..4}
    x first {6,15,24} frames of a 40x32x24 stack, through the autotuned, graph-captured
    fused sweep, against initialise/calculateSRRF/readSRRFbuffer per configuration."""
    st = stack(24, 40, 32, seed=303)
    sw = pkg.ParametersSweep(40, 32, C3_MAGS, C3_FWHM, C3_SENS, C3_NFS, autotune=True,
                             use_graph=True)
    sw.run(st)
    shapes = set()
    for im, M in enumerate(C3_MAGS):
        for i_f, fwhm in enumerate(C3_FWHM):
            sh = sw.shape(im, i_f)
            shapes.add((sh["nw"], sh["taps_in_lds"], sh["frames_per_pass"]))
            for iq, nf in enumerate(C3_NFS):
                for i_s, S in enumerate(C3_SENS):
                    exp = plain(pkg, st, M, fwhm, S, nf)
                    assert_parity(sw.read(im, iq, i_s, i_f), exp,
                                  f"M{M} nf{nf} S{S} fwhm{fwhm}")
    sw.close()
    print("autotuned shapes:", shapes)


@pytest.mark.parametrize("M", C3_MAGS)
@pytest.mark.parametrize("fwhm", C3_FWHM)
def test_every_runnable_launch_shape_is_bit_identical(pkg, M, fwhm):
    """Every launch shape the autotuner may pick (4 to 16 waves, tap tables in LDS or
    through scalar loads, one, two or four frames per pass) gives the default shape's bits
    for this (magnification, fwhm), on a 40x32 stack in two loads."""
    st = stack(24, 40, 32, seed=404 + M)
    ref, shape0 = None, None
    ran = []
    for nw in ([M * M] if M * M <= 4 else []) + [4, 8, 12, 16]:
        for tl in (1, 0):
            for fp in (1, 2, 4):
                e = pkg.LiveSRRF_CL()
                e.checkDevices()
                e.initialise(40, 32, M, fwhm, 2.0, 12, 24, 20000, e.allCLdevices[0], True, True,
                             False, None, 0, 1)
                if shape0 is None:
                    shape0 = e.launchShape()
                try:
                    e.setLaunchShape(nw=nw, taps_in_lds=tl, frames_per_pass=fp)
                except pkg._capi.EsrrfError:
                    e.release()
                    continue      # not runnable for this geometry (LDS budget)
                e.resetFramePosition()
                e.loadDriftXYGPUbuffer(np.zeros(24), np.zeros(24))
                e.calculateSRRF(st[:12])
                e.calculateSRRF(st[12:])
                out = np.asarray(e.readSRRFbuffer()).copy()
                e.release()
                ran.append((nw, tl, fp))
                if ref is None:
                    e0 = pkg.LiveSRRF_CL()
                    e0.checkDevices()
                    e0.initialise(40, 32, M, fwhm, 2.0, 12, 24, 20000, e0.allCLdevices[0],
                                  True, True, False, None, 0, 1)
                    e0.resetFramePosition()
                    e0.loadDriftXYGPUbuffer(np.zeros(24), np.zeros(24))
                    e0.calculateSRRF(st[:12])
                    e0.calculateSRRF(st[12:])
                    ref = np.asarray(e0.readSRRFbuffer()).copy()
                    e0.release()
                assert np.array_equal(out.view(np.uint32), ref.view(np.uint32)), (nw, tl, fp)
    assert len(ran) >= 4, ran
    # the default shape is among the runnable ones
    assert (shape0["nw"], shape0["taps_in_lds"], shape0["frames_per_pass"]) in ran


@pytest.mark.parametrize("fwhms", [[1.0, 1.5, 2.0, 2.5, 3.0], [2.2, 0.9, 1.4, 3.0, 1.1, 2.7, 1.8]])
def test_multi_fwhm_emit_bit_identical_to_per_fwhm_emit(pkg, monkeypatch, fwhms):
    """One emit launch per magnification and group of up to five fwhm values (one angular
    term per sample, a weight per fwhm, 0 outside its gate) gives every configuration the
    same bits as one emit launch per fwhm; seven unsorted fwhm values make a group of five
    and one of two."""
    st = stack(14, 36, 28, seed=515)
    mags, sens, nfs = [2, 5], [1.0, 2.5], [6, 14]
    res = {}
    for per in ("1", "0"):
        monkeypatch.setenv("ESRRF_SWEEP_PER_FWHM", per)
        sw = pkg.ParametersSweep(36, 28, mags, fwhms, sens, nfs, autotune=(per == "0"),
                                 use_graph=True)
        sw.run(st)
        res[per] = {(im, iq, i_s, i_f): np.asarray(sw.read(im, iq, i_s, i_f)).copy()
                    for im in range(len(mags)) for iq in range(len(nfs))
                    for i_s in range(len(sens)) for i_f in range(len(fwhms))}
        sw.close()
    for k, v in res["1"].items():
        assert np.array_equal(v.view(np.uint32), res["0"][k].view(np.uint32)), k


def plain_fht(pkg, st, M, fwhm, S, nf):
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(st.shape[2], st.shape[1], M, fwhm, S, nf, nf, 20000, e.allCLdevices[0], True,
                 True, True, None, 0, 1)
    e.resetFramePosition()
    e.loadDriftXYGPUbuffer(np.zeros(nf), np.zeros(nf))
    e.calculateSRRF(st[:nf])
    out = np.asarray(e.readSRRFbuffer()).copy()
    e.release()
    return out


@pytest.mark.parametrize("use_graph", [False, True])
def test_fht_sweep_bit_identical_to_per_configuration_engine(pkg, use_graph):
    """The sweep in FHT mode (the reference sweep's default, gui/ParametersSweep_.java:117,
    452): frames upsampled once per magnification, one CGLH emit launch per fwhm, one FHT
    accumulation per fwhm for every (frame count, sensitivity); each configuration equals
    initialise(doFHTinterpolation) + calculateSRRF + readSRRFbuffer bit for bit (no MP
    correction in FHT mode, LiveSRRF_CL.java:762)."""
    st = stack(10, 24, 20, seed=616)
    mags, fwhms, sens, nfs = [2, 3], [1.0, 1.5, 2.0], [1.0, 2.5], [5, 10]
    sw = pkg.ParametersSweep(24, 20, mags, fwhms, sens, nfs, use_graph=use_graph, fht=True)
    sw.run(st)
    for im, M in enumerate(mags):
        for iq, nf in enumerate(nfs):
            for i_s, S in enumerate(sens):
                for i_f, fwhm in enumerate(fwhms):
                    got = np.asarray(sw.read(im, iq, i_s, i_f))
                    exp = plain_fht(pkg, st, M, fwhm, S, nf)
                    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), \
                        (M, nf, S, fwhm)
    with pytest.raises(pkg._capi.EsrrfError):
        sw.shape(0, 0)                # FHT mode has no tiled 2-D launch shape
    sw.close()
