"""Write the golden vectors of tests/golden/: the outputs of the reference's own OpenCL
kernels (resources/liveSRRF.cl, liveSRRF_noInterpolation.cl, live3DSRRF.cl, built from their
sources into oracle/_ref/ by oracle/ref/build_ref.py) for small seeded inputs, run through
the AMD OpenCL runtime with LiveSRRF_CL's enqueue order (oracle/refrun.py).

Needs an MI355X and a built oracle/_ref/:

    python tests/golden/make_ref_vectors.py [OUT_DIR]

(OUT_DIR defaults to tests/golden/.)

One ref_<config>.npz per configuration holds the input stack exactly as the kernels saw it
(float32; FHT: the upsampled frames and the raw frames; 3-D: the reshaped planes), the
drift (BIC), the result [4][nvox] (NaN -> 0), the MP map and, for BIC, the interpolated
gradients of frame 0.  The CPU tests compare the oracle with them.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)


def main():
    out_dir = sys.argv[1] if len(sys.argv) > 1 else HERE
    os.makedirs(out_dir, exist_ok=True)
    import __graft_entry__
    __graft_entry__.load()
    from nanoj_esrrf_amd import synth
    from oracle import fht, refrun
    if not refrun.available():
        raise SystemExit("oracle/_ref is not built")
    for name, cfg in refrun.configs().items():
        mode, w, h, nP, M, fwhm, S, nF, iw = cfg
        if name == "bic_plain_64x64_m5":
            continue     # large; covered by the GPU tests
        rec = {"config": np.array(cfg[1:], dtype=np.float64), "mode": np.array(mode)}
        if mode == "bic":
            st = synth.blinking_stack(nF, w, h, max(20, w * h // 14), 4, seed=700 + nF + w)
            st = st.astype(np.float32)
            loads = [4, 4, 3] if name.endswith("loads") else [nF]
            rng = np.random.default_rng(w * h)
            dx = rng.uniform(-0.2, 0.2, nF).astype(np.float32) if name == "bic_40x24_m2" \
                else np.zeros(nF, np.float32)
            dy = rng.uniform(-0.2, 0.2, nF).astype(np.float32) if name == "bic_40x24_m2" \
                else np.zeros(nF, np.float32)
            drift = np.concatenate([np.float32(M) * dx, np.float32(M) * dy]).astype(np.float32)
            out, mp, gint = refrun.run(name, st, loads=loads, drift=drift, want_gint=True)
            rec.update(frames=st, loads=np.array(loads), drift_x=dx, drift_y=dy, out=out, mp=mp,
                       gint=gint)
        elif mode == "fht":
            st = synth.blinking_stack(nF, w, h, max(20, w * h // 14), 4, seed=800 + nF + w)
            st = st.astype(np.float32)
            up = fht.fht_stack(st, M, True)
            out, _, _ = refrun.run(name, up)
            rec.update(frames=st, upsampled=up, out=out)
        else:
            planes = synth.blinking_stack_3d(nF, w, h, nP, n_emitters=60, n_filaments=4,
                                             seed=900 + nF + w).astype(np.float32)
            z = np.zeros(nP)
            out, mp, _ = refrun.run(name, planes, st3d=refrun.st3d_of(z, z, z))
            rec.update(planes=planes, out=out, mp=mp)
        path = os.path.join(out_dir, f"ref_{name}.npz")
        np.savez_compressed(path, **rec)
        print(f"wrote {path} ({os.path.getsize(path)} bytes)", flush=True)


if __name__ == "__main__":
    main()
