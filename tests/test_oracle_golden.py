"""The CPU oracle against golden vectors of the reference's own OpenCL kernels
(tests/golden/ref_*.npz, written by tests/golden/make_ref_vectors.py from the
resources/*.cl sources run through the AMD OpenCL runtime in LiveSRRF_CL's enqueue order).
CPU only.

Tolerance: per channel, normalised by the reference image's max |value|,
|a - b| <= 1e-5 + 1e-4 |b|; MP maps rtol 1e-4; integer-valued gradients exact.
"""
import glob
import os

import numpy as np
import pytest

from conftest import REPO, normalised_close

FILES = sorted(glob.glob(os.path.join(REPO, "tests", "golden", "ref_*.npz")))


def assert_close(got, exp, what):
    got = np.asarray(got).reshape(4, -1)
    exp = np.asarray(exp).reshape(4, -1)
    for c, name in enumerate(("AVG", "VAR", "TAC2", "INT")):
        ok, ex, idx = normalised_close(got[c], exp[c])
        assert ok, f"{what}: {name} differs (excess {ex:.3g} at {idx})"


def test_golden_vectors_present():
    assert len(FILES) >= 9


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f)[4:-4] for f in FILES])
def test_oracle_matches_reference_kernel_vectors(O, pkg, path):
    d = np.load(path, allow_pickle=False)
    w, h, nP, M, fwhm, S, nF, iw = [float(v) for v in d["config"]]
    w, h, nP, M, nF, iw = int(w), int(h), int(nP), int(M), int(nF), bool(iw)
    mode = str(d["mode"])
    name = os.path.basename(path)
    if mode == "bic":
        loads = [int(v) for v in d["loads"]]
        o = O.OracleLiveSRRF(w, h, M, fwhm, S, max(loads), nF, iw, True, False)
        o.loadDriftXYGPUbuffer(d["drift_x"], d["drift_y"])
        a = 0
        for n in loads:
            o.calculateSRRF(d["frames"][a:a + n])
            if a == 0:
                g = o.readGradientBuffers()
                assert np.array_equal(g.reshape(-1), d["gint"]), f"{name}: gradients"
            a += n
        assert_close(o.readSRRFbuffer(), d["out"], name)
        np.testing.assert_allclose(o.mp, d["mp"], rtol=1e-4)
    elif mode == "fht":
        o = O.OracleLiveSRRF(w, h, M, fwhm, S, nF, nF, iw, True, True)
        o.calculateSRRF(d["frames"])
        assert_close(o.readSRRFbuffer(), d["out"], name)
    else:
        from nanoj_esrrf_amd import synth
        z = np.zeros(nP)
        calib = O.Calib3D(z, z, z, np.arange(nP, dtype=np.float64), np.ones(nP), 0.0)
        o = O.OracleLiveSRRF(3 * w, 3 * h, M, fwhm, S, nF, nF, iw, True, False, calib3d=calib)
        o.calculateSRRF(synth.montage3x3(d["planes"]))
        assert_close(o.readSRRFbuffer(), d["out"], name)
        np.testing.assert_allclose(o.mp, d["mp"], rtol=1e-4)
