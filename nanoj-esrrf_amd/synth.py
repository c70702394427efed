"""Seeded synthetic blinking-emitter stacks for the eSRRF workloads.

Emitters sit on random quadratic Bezier filaments (uniform in the curve parameter); each
blinks as a two-state Markov chain started from its steady state; the PSF is a pixel-
integrated Gaussian; every frame is Poisson(signal + background) plus Gaussian read noise,
rounded and clipped to uint16.  This is the project's own generator: the reference's
simulators (gui/FluorescenceSimulator_.java) are ImageJ plugins and are not reused.
"""
from __future__ import annotations

import math

import numpy as np

SEED_BASE = 20261203


def _bezier_points(rng, n_fil, n_em, lo, hi, dim=2):
    ctrl = rng.uniform(lo, hi, size=(n_fil, 3, dim))
    fil = np.arange(n_em) % n_fil
    t = rng.uniform(0.0, 1.0, size=n_em)[:, None]
    p0, p1, p2 = ctrl[fil, 0], ctrl[fil, 1], ctrl[fil, 2]
    return (1 - t) ** 2 * p0 + 2 * (1 - t) * t * p1 + t ** 2 * p2


def _pixel_integrated_profile(centres, n, sigma):
    """(n_em, n) matrix: integral of a unit Gaussian over each pixel [k, k+1)."""
    from scipy.special import erf
    edges = np.arange(n + 1, dtype=np.float64)[None, :]
    z = (edges - centres[:, None]) / (math.sqrt(2.0) * sigma)
    c = 0.5 * erf(z)
    return np.diff(c, axis=1)


def _blink_states(rng, n_frames, n_em, p_on, p_off):
    state = rng.uniform(size=n_em) < p_on / (p_on + p_off)
    out = np.empty((n_frames, n_em), dtype=bool)
    for f in range(n_frames):
        out[f] = state
        u = rng.uniform(size=n_em)
        state = np.where(state, u >= p_off, u < p_on)
    return out


def blinking_stack(n_frames=100, width=64, height=64, n_emitters=300, n_filaments=8,
                   psf_sigma=1.3, p_on=0.05, p_off=0.5, photons=500.0, background=100.0,
                   read_noise=2.0, seed=SEED_BASE + 1):
    """2-D stack [n_frames][height][width] uint16."""
    rng = np.random.default_rng(seed)
    pos = _bezier_points(rng, n_filaments, n_emitters, 2.0, min(width, height) - 2.0)
    pos[:, 0] = np.clip(pos[:, 0] * (width / min(width, height)), 0, width)
    pos[:, 1] = np.clip(pos[:, 1] * (height / min(width, height)), 0, height)
    px = _pixel_integrated_profile(pos[:, 0], width, psf_sigma)     # (E, W)
    py = _pixel_integrated_profile(pos[:, 1], height, psf_sigma)    # (E, H)
    states = _blink_states(rng, n_frames, n_emitters, p_on, p_off)
    out = np.empty((n_frames, height, width), dtype=np.uint16)
    for f in range(n_frames):
        on = np.nonzero(states[f])[0]
        sig = (py[on].T * photons) @ px[on] if on.size else np.zeros((height, width))
        lam = sig + background
        img = rng.poisson(lam).astype(np.float64) + rng.normal(0.0, read_noise, size=lam.shape)
        out[f] = np.clip(np.rint(img), 0, 65535).astype(np.uint16)
    return out


def blinking_stack_3d(n_frames=512, width=128, height=128, n_planes=9, depth_planes=None,
                      n_emitters=1500, n_filaments=30, psf_sigma=1.3, psf_sigma_z=2.0,
                      p_on=0.05, p_off=0.5, photons=500.0, background=100.0, read_noise=2.0,
                      seed=SEED_BASE + 5):
    """3-D stack [n_frames][n_planes][height][width] uint16 (planes 0.3 um apart;
    the emitters fill a volume of `depth_planes` planes, default n_planes)."""
    rng = np.random.default_rng(seed)
    depth = float(n_planes if depth_planes is None else depth_planes)
    pos = _bezier_points(rng, n_filaments, n_emitters, 0.0, 1.0, dim=3)
    pos[:, 0] = 2.0 + pos[:, 0] * (width - 4.0)
    pos[:, 1] = 2.0 + pos[:, 1] * (height - 4.0)
    pos[:, 2] = -0.5 + pos[:, 2] * depth
    px = _pixel_integrated_profile(pos[:, 0], width, psf_sigma)
    py = _pixel_integrated_profile(pos[:, 1], height, psf_sigma)
    pz = np.exp(-0.5 * ((np.arange(n_planes)[None, :] - pos[:, 2:3]) / psf_sigma_z) ** 2)
    states = _blink_states(rng, n_frames, n_emitters, p_on, p_off)
    out = np.empty((n_frames, n_planes, height, width), dtype=np.uint16)
    for f in range(n_frames):
        on = np.nonzero(states[f])[0]
        for z in range(n_planes):
            a = photons * pz[on, z]
            sig = (py[on].T * a) @ px[on] if on.size else np.zeros((height, width))
            lam = sig + background
            img = rng.poisson(lam).astype(np.float64) + rng.normal(0.0, read_noise, size=lam.shape)
            out[f, z] = np.clip(np.rint(img), 0, 65535).astype(np.uint16)
    return out


def montage3x3(planes: np.ndarray) -> np.ndarray:
    """Tile [n][9][h][w] planes into the 3x3 multi-focus montage [n][3h][3w] that
    LiveSRRF_CL.reshapeImageStack3D splits (plane z at cell roi=z)."""
    n, nP, h, w = planes.shape
    out = np.zeros((n, 3 * h, 3 * w), dtype=planes.dtype)
    for z in range(nP):
        i, j = z % 3, z // 3
        out[:, j * h:(j + 1) * h, i * w:(i + 1) * w] = planes[:, z]
    return out


# Workloads named by what they are.
def plain_64x64x100(seed=SEED_BASE + 1):
    return blinking_stack(100, 64, 64, 300, 8, background=100.0, seed=seed)


def production_256x256x1000(seed=SEED_BASE + 2, n_frames=1000):
    return blinking_stack(n_frames, 256, 256, 2000, 40, background=200.0, seed=seed)


def sweep_256x256x500(seed=SEED_BASE + 3):
    """Parameter-sweep workload: 500 frames of 256x256, the production emitter density."""
    return blinking_stack(500, 256, 256, 2000, 40, background=200.0, seed=seed)


def rse_128x128x500(seed=SEED_BASE + 4):
    """Parameter-selection workload: 500 frames of 128x128 at the production density
    (a quarter of the field: 500 emitters on 10 filaments)."""
    return blinking_stack(500, 128, 128, 500, 10, background=200.0, seed=seed)


def mfm_9x128x128x512(seed=SEED_BASE + 5):
    """3-D workload: 512 frames of 9 focal planes of 128x128 (1500 emitters on 30 filaments
    in a 9-plane volume) as planes [n][9][128][128]."""
    return blinking_stack_3d(512, 128, 128, 9, seed=seed)
