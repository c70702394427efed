"""Rolling analysis (gui/LiveSRRF_optimised_.java:238-384): every SR frame of the fused
rolling engine (one RGC pass per frame shared by overlapping windows) against the plugin's
loop of the engine per window (initialise, reset, calculateSRRF over the window's frames,
readSRRFbuffer).  Both apply the same per-frame CGLH and intensity and the same temporal
accumulation in the same order, so the results are identical bit for bit; the parity
tolerance of the engine is asserted as well.
"""
import zipfile

import numpy as np
import pytest

from conftest import hip_runtime

from conftest import normalised_close

pytestmark = pytest.mark.gpu


def stack(n, w, h, seed):
    from nanoj_esrrf_amd import synth
    return synth.blinking_stack(n, w, h, max(20, w * h // 14), 4, seed=seed)


def per_window(pkg, st, M, fwhm, S, nF, gap, intw=True, mp=True):
    plan = pkg._capi.plan_rolling(st.shape[0], nF, gap, True, st.shape[0])
    nF, gap, nSR = plan["nFrameForSRRF"], plan["gap"], plan["nSRRFframe"]
    e = pkg.LiveSRRF_CL()
    e.checkDevices()
    e.initialise(st.shape[2], st.shape[1], M, fwhm, S, nF, nF, 20000, e.allCLdevices[0], intw,
                 mp, False, None, 0, 1)
    outs = []
    for r in range(nSR):
        e.resetFramePosition()
        e.loadDriftXYGPUbuffer(np.zeros(nF), np.zeros(nF))
        e.calculateSRRF(st[r * gap:r * gap + nF])
        outs.append(np.asarray(e.readSRRFbuffer()).copy())
    e.release()
    return outs


@pytest.mark.parametrize("n,w,h,M,fwhm,S,nF,gap,chunk,intw,mp", [
    (40, 24, 20, 3, 1.5, 2, 12, 5, 7, True, True),      # overlapping windows, small chunks
    (37, 20, 22, 2, 1.0, 1, 10, 10, 16, True, True),    # disjoint (no rolling), ragged end
    (45, 18, 16, 4, 2.0, 3, 8, 13, 6, False, False),    # frame gap > window: skipped frames
    (30, 16, 16, 5, 1.5, 2, 30, 0, 256, True, True),    # one window over the whole stack
    (64, 16, 20, 2, 1.5, 2, 20, 1, 9, True, True),      # a window at every frame
    (48, 20, 18, 3, 1.5, 2, 16, 4, 10, True, False),    # float frames with negative values
    (40, 16, 16, 2, 1.0, 1.5, 10, 3, 8, True, True),    # a non-integer sensitivity (powf)
])
def test_rolling_matches_engine_per_window(pkg, n, w, h, M, fwhm, S, nF, gap, chunk, intw, mp):
    st = stack(n, w, h, seed=n + w)
    if n == 48:
        # intensities below zero: v C^S is -0 where the CGLH is 0, including at the first
        # frame of windows, whose sums then start from a negative zero
        st = st.astype(np.float32) - np.float32(st.mean())
    exp = per_window(pkg, st, M, fwhm, S, nF, gap, intw, mp)
    r = pkg.Rolling(w, h, M, fwhm, S, nF, gap, intw, mp, chunk_frames=chunk)
    nsr = r.begin(st)
    assert nsr == len(exp)
    for k in range(nsr):
        idx, got = r.next()
        assert idx == k
        for c in range(4):
            ok, ex, at = normalised_close(got[c], exp[k][c])
            assert ok, f"SR frame {k} channel {c}: excess {ex:.3g} at {at}"
        assert np.array_equal(got, exp[k]), f"SR frame {k}: not bit-identical"
    with pytest.raises(pkg._capi.EsrrfError):
        r.next()
    r.close()


def test_rolling_device_stack_and_reuse(pkg):
    import ctypes
    hip = hip_runtime()
    st = stack(33, 20, 20, seed=5)
    exp = per_window(pkg, st, 3, 1.5, 2, 11, 4)
    r = pkg.Rolling(20, 20, 3, 1.5, 2, 11, 4, chunk_frames=10)
    d_st, d_out = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d_st), st.nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(d_out), 4 * 60 * 60 * 4) == 0
    assert hip.hipMemcpy(d_st, st.ctypes.data, st.nbytes, 1) == 0          # host to device
    out = np.empty((4, 60, 60), np.float32)
    for _ in range(2):   # a second run on the same object starts over
        n = r.begin_device(d_st.value, pkg._capi.ESRRF_U16, st.shape[0])
        for k in range(n):
            assert r.next_device(d_out.value) == k
            assert hip.hipDeviceSynchronize() == 0
            assert hip.hipMemcpy(out.ctypes.data, d_out, out.nbytes, 2) == 0   # device to host
            assert np.array_equal(out, exp[k])
    r.close()
    hip.hipFree(d_st)
    hip.hipFree(d_out)


def test_driver_writes_the_plugins_zip_and_falls_back_for_drift(pkg, tmp_path):
    from nanoj_esrrf_amd import rolling
    st = stack(26, 16, 18, seed=8)
    drv = pkg.LiveSRRF_optimised(magnification=2, fwhm=1.5, sensitivity=2, nFrameForSRRF=10,
                                 doRollingAnalysis=True, frameGap=4, calculateVAR=True,
                                 getInterpolatedImage=True, chunk_frames=8)
    stacks = drv.run(st)
    assert drv.last_path == "fused" and set(stacks) == {"AVG", "VAR", "INT"}
    exp = per_window(pkg, st, 2, 1.5, 2, 10, 4)
    assert stacks["AVG"].shape == (len(exp), 36, 32)
    for k in range(len(exp)):
        assert np.array_equal(stacks["VAR"][k], exp[k][1])
    path = drv.run(st, writeToDisk=str(tmp_path / "out.zip"))
    with zipfile.ZipFile(path) as z:
        names = z.namelist()
        assert len(names) == 3 * len(exp) and "eSRRF (AVG) - frame=1.tif" in names
        img = rolling.read_tiff(z.read(f"eSRRF (INT) - frame={len(exp)}.tif"))
        assert np.array_equal(img, exp[-1][3])
    # zero drift through the engine loop gives the same images
    drift_stacks = drv.run(st, drift=lambda r: (np.zeros(10, np.float32),) * 2)
    assert drv.last_path == "engine"
    for k in range(len(exp)):
        assert np.array_equal(drift_stacks["AVG"][k], stacks["AVG"][k])
