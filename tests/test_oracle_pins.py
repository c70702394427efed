"""Pins of the CPU oracle (oracle/esrrf_oracle.c, oracle/fht.py, oracle/oracle.py).

The reference ships no golden vectors, known-answer tests or fixtures for this path
(HelperUtils.compareResults expects a testdat/ folder that does not exist,
HelperUtils.java:94).  Each check below is fixed by mathematics the reference's code
implies -- closed forms, invariants and special cases -- and is chosen so that a plausible
restatement mistake (wrong sign, dropped term, transposed or shifted operand, wrong
normalisation) fails it.  The index conventions of the RGC reads are additionally pinned
against the reference kernels themselves on the GPU (tests/test_gpu_ref.py).
"""
import math

import numpy as np
import pytest

f32 = np.float32


# ---------------------------------------------------------------------------------------
# Constants (LiveSRRF_CL.java:304-305,317-320; :491-500) and Appendix-B values
# ---------------------------------------------------------------------------------------
# tSO: `"" + 2 * sigma + 1` (LiveSRRF_CL.java:318) is Java string concatenation, so the
# kernel text is Float.toString(2*sigma) followed by "1"; the texts below are typed from
# Java's rules (JLS 15.18.1, Float.toString), not computed by the code under test.
@pytest.mark.parametrize("fwhm,sigma,radius,tso_text,tSS", [
    (1.0, 0.424809, 1.5, "0.849617661", 0.36093), (1.5, 0.637213, 2.0, "1.27442651", 0.81208),
    (2.0, 0.849618, 2.5, "1.69923531", 1.44370), (2.5, 1.062022, 3.0, "2.12404421", 2.25578),
    (3.0, 1.274426, 3.5, "2.5488531", 3.24833)])
def test_bic_constants(O, pkg, fwhm, sigma, radius, tso_text, tSS):
    k = O.bic_constants(fwhm)
    s = f32(fwhm) / f32(2.354)
    assert k["sigma"] == float(s)
    assert k["radius"] == float(f32(int(f32(4.0) * s)) / f32(2) + f32(1))
    assert k["tSS"] == float(f32(f32(2) * s) * s)
    assert abs(k["sigma"] - sigma) < 1e-6 and k["radius"] == radius
    assert abs(k["tSS"] - tSS) < 1e-5
    # (float) of the double literal, as `(float) tSO` (liveSRRF.cl:291) reads it
    assert k["tSO"] == float(f32(float(tso_text)))
    assert O.java_float_to_string(f32(2) * s) + "1" == tso_text
    assert pkg._capi.bic_tso_text(fwhm) == tso_text
    assert pkg._capi.bic_constants(fwhm)["tSO"] == k["tSO"]


# Float.toString cases fixed by the Java SE specification (decimal form for 1e-3 <= |x| <
# 1e7, computerized scientific notation otherwise, one- and two-digit candidates when one
# digit suffices, at least one fraction digit).
@pytest.mark.parametrize("x,text", [
    (2.0, "2.0"), (100.0, "100.0"), (0.1, "0.1"), (0.5, "0.5"), (1e-3, "0.001"),
    (1e-4, "1.0E-4"), (1e7, "1.0E7"), (9999999.0, "9999999.0"), (1234567.8, "1234567.8"),
    (1.4e-45, "1.4E-45"), (3.4028235e38, "3.4028235E38"), (1.17549435e-38, "1.1754944E-38"),
    (-2.5e-5, "-2.5E-5"), (0.0, "0.0"), (-0.0, "-0.0"), (1.0 / 3.0, "0.33333334")])
def test_java_float_to_string_cases(O, x, text):
    assert O.java_float_to_string(f32(x)) == text


def _java_text_numpy(x):
    """Float.toString through numpy's Dragon4 shortest digits -- an implementation
    independent of the oracle's and the engine's -- with Java's layout and its rule for
    one-digit decimals."""
    from fractions import Fraction
    x = f32(x)
    sci = np.format_float_scientific(x, unique=True, trim="-", exp_digits=1)
    if len(sci.split("e")[0].lstrip("-").replace(".", "")) == 1:
        two = np.format_float_scientific(x, precision=1, unique=False, exp_digits=1)
        if f32(two) == x and abs(Fraction(two) - Fraction(float(x))) < \
                abs(Fraction(sci) - Fraction(float(x))):
            sci = two
    mant, ex = sci.split("e")
    sign = "-" if mant.startswith("-") else ""
    dg = mant.lstrip("-").replace(".", "")
    ex = int(ex)
    if -3 <= ex <= 6:
        if ex >= 0:
            return f"{sign}{dg[:ex + 1].ljust(ex + 1, '0')}.{dg[ex + 1:] or '0'}"
        return f"{sign}0.{'0' * (-ex - 1)}{dg}"
    return f"{sign}{dg[0]}.{dg[1:] or '0'}E{ex}"


def test_java_float_to_string_matches_dragon4(O, pkg):
    rng = np.random.default_rng(5)
    xs = np.concatenate([
        (f32(2) * (rng.uniform(0.3, 12.0, 3000).astype(f32) / f32(2.354))).astype(f32),
        f32(10.0) ** rng.uniform(-44, 38, 3000).astype(f32),
        np.array([f32(2.0) ** k for k in range(-149, 128)], f32)])
    for x in xs:
        if not np.isfinite(x) or x == 0:
            continue
        assert O.java_float_to_string(x) == _java_text_numpy(x), x
    # the engine's restatement, through the fwhm it is asked about
    for fwhm in np.round(np.arange(0.5, 6.0, 0.01), 2):
        s = f32(fwhm) / f32(2.354)
        txt = _java_text_numpy(f32(2) * s) + "1"
        assert pkg._capi.bic_tso_text(float(fwhm)) == txt, fwhm
        assert pkg._capi.bic_constants(float(fwhm))["tSO"] == float(f32(float(txt))), fwhm


def test_bic_gated_samples_production_geometry(O):
    """The production geometry (256x256, M=5, fwhm 1.5): with tSO = 1.27442651 the gate of
    liveSRRF.cl:291 keeps about 20.3 samples per output pixel (3.33e7 per frame), against
    64.6 for a gate of 2*sigma+1 -- counted by brute force over the reference's window."""
    k = O.bic_constants(1.5)
    M, w, h = 5, 256, 256
    n = 0
    r2 = int(f32(2) * f32(k["radius"]))
    r2hi = int(f32(2) * f32(k["radius"]) + f32(1))
    # phase geometry of interior pixels (every vx - xc is exact for X >= 8)
    for sy in range(M):
        for sx in range(M):
            xM, yM = 16 * M + sx, 16 * M + sy
            xc = f32((xM + 0.5) / M)
            yc = f32((yM + 0.5) / M)
            bx, by = int(f32(2) * xc), int(f32(2) * yc)
            for j in range(-r2, r2hi + 1):
                vy = (f32(by) + f32(j)) / f32(2)
                for i in range(-r2, r2hi + 1):
                    vx = (f32(bx) + f32(i)) / f32(2)
                    dx, dy = f32(vx - xc), f32(vy - yc)
                    d = np.sqrt(f32(f32(dx * dx) + f32(dy * dy)))
                    if d != 0 and d <= f32(k["tSO"]):
                        n += 1
    per_px = n / (M * M)
    assert 20.0 < per_px < 20.6, per_px
    assert abs(per_px * w * h * M * M - 3.33e7) < 0.01e7


def test_fht_constants(O):
    k = O.fht_constants(1.5, 5)
    assert k["radius"] == 10.0
    assert abs(k["tSO"] - 11.37213) < 1e-4 and abs(k["tSS"] - 20.30204) < 1e-4
    k = O.fht_constants(3.0, 5)
    assert k["radius"] == 15.0


def test_product_constants_match_oracle(O, pkg):
    for fwhm in (0.8, 1.0, 1.37, 1.5, 2.0, 2.5, 3.0, 4.2):
        a, b = O.bic_constants(fwhm), pkg._capi.bic_constants(fwhm)
        assert a == b
        for M in (1, 2, 5, 8):
            assert O.fht_constants(fwhm, M) == pkg._capi.fht_constants(fwhm, M)


# ---------------------------------------------------------------------------------------
# Gradients (liveSRRF.cl:187-214; liveSRRF_noInterpolation.cl:138-159)
# ---------------------------------------------------------------------------------------
def test_two_point_gradient_exact(O):
    rng = np.random.default_rng(0)
    px = rng.integers(0, 4000, size=(3, 9, 13)).astype(np.float32)
    gx, gy = O.gradient_2point(px)
    ex = np.diff(px, axis=2, prepend=px[:, :, :1])     # backward difference, edge 0
    ey = np.diff(px, axis=1, prepend=px[:, :1, :])
    assert np.array_equal(gx, ex) and np.array_equal(gy, ey)


def test_robx_gradient_linear_ramp(O):
    y, x = np.mgrid[0:10, 0:12].astype(np.float32)
    px = (3.0 * x - 2.0 * y + 5.0)[None].astype(np.float32)
    gx, gy = O.gradient_robx(px)
    # interior: Gx = 2*alpha, Gy = 2*beta for f = alpha x + beta y (rotated Roberts cross)
    assert np.all(gx[0, 1:, 1:] == 6.0) and np.all(gy[0, 1:, 1:] == -4.0)
    # first column: x0 = x1 -> Gx = p01 - p00 + ... collapses to the vertical terms
    assert np.all(gx[0, 1:, 0] == 0.0) and np.all(gy[0, 1:, 0] == -4.0)


# ---------------------------------------------------------------------------------------
# Catmull-Rom kernel and interpolation (liveSRRF.cl:26-133)
# ---------------------------------------------------------------------------------------
def test_cubic_kernel_values(O):
    assert O.cubic(0.0) == 1.0
    for v in (1.0, -1.0, 2.0, -2.0, 2.5, 7.0):
        assert O.cubic(v) == 0.0
    assert O.cubic(0.5) == 0.5625 and O.cubic(-0.5) == 0.5625
    assert O.cubic(1.5) == -0.0625 and O.cubic(-1.5) == -0.0625
    for t in np.linspace(0, 1, 17):
        s = sum(O.cubic(float(t) - k) for k in (-1, 0, 1, 2))
        assert abs(s - 1.0) < 1e-6      # partition of unity


def test_interp_reproduces_quadratics_inside_and_lines_at_borders(O):
    h, w = 11, 13
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    quad = (0.25 * x * x - 0.5 * x * y + 0.125 * y * y + 2 * x - y + 3).astype(np.float32)
    lin = (1.5 * x - 0.75 * y + 4.0).astype(np.float32)
    rng = np.random.default_rng(1)
    for _ in range(60):
        xs, ys = rng.uniform(1.05, w - 2.05), rng.uniform(1.05, h - 2.05)
        exact = 0.25 * xs * xs - 0.5 * xs * ys + 0.125 * ys * ys + 2 * xs - ys + 3
        assert abs(O.interp(quad, xs, ys) - exact) < 2e-4 * max(1.0, abs(exact))
    # border regime: bilinear with extrapolation reproduces lines everywhere, also outside
    for xs, ys in [(-0.7, 3.2), (0.3, 0.2), (w - 1.2, 5.5), (w + 0.4, h + 0.3), (4.5, -1.25)]:
        exact = 1.5 * xs - 0.75 * ys + 4.0
        assert abs(O.interp(lin, xs, ys) - exact) < 1e-4


def test_gradient_interp_even_points_copy_and_half_points(O):
    rng = np.random.default_rng(2)
    g = rng.integers(-500, 500, size=(1, 10, 12)).astype(np.float32)
    gi = O.gradient_interp(g)
    # even points of the interior are exact copies (cubic(0)=1, cubic(+-1)=cubic(+-2)=0)
    assert np.array_equal(gi[0, 2:16:2, 2:20:2], g[0, 1:8, 1:10])
    # an interior half point is the (-1, 9, 9, -1)/16 stencil
    y, x = 9, 11                     # (x/2, y/2) = (5.5, 4.5)
    w1 = np.array([-1, 9, 9, -1]) / 16.0
    exp = w1 @ g[0, 3:7, 4:8].astype(np.float64) @ w1
    assert gi[0, y, x] == np.float32(exp)


# ---------------------------------------------------------------------------------------
# RGC (liveSRRF.cl:238-401) on synthetic gradient fields
# ---------------------------------------------------------------------------------------
def _bic_frame(O, gxi, gyi, px, M, fwhm, S, intW=1, n=2):
    import ctypes
    k = O.bic_constants(fwhm)
    h, w = px.shape
    cfg = O._BicCfg(w, h, M, S, k["radius"], k["tSS"], k["tSO"], n, intW)
    out = np.zeros(4 * w * M * h * M, np.float32)
    prev = np.zeros(w * M * h * M, np.float32)
    drift = np.zeros(2 * n, np.float32)
    O.lib().orc_bic_rgc_frame(ctypes.byref(cfg), np.ascontiguousarray(px[None]),
                              np.ascontiguousarray(gxi[None]), np.ascontiguousarray(gyi[None]),
                              drift, 0, 0, out, prev, 1)
    return out.reshape(4, h * M, w * M), prev.reshape(h * M, w * M)


@pytest.mark.parametrize("M,xM,yM,fwhm", [(5, 52, 47, 1.5), (4, 40, 37, 2.5), (2, 21, 17, 1.0)])
def test_rgc_convergent_field_gives_one(O, M, xM, yM, fwhm):
    """A gradient field that points exactly at the pixel centre from every sample position
    (G(v) = -(v - c), read at GxInt[m-1][n], GyInt[m][n-1] for sample (n/2, m/2)) gives
    CGLH = 1 there; the outward field gives 0."""
    h = w = 24
    xc = np.float32((xM + 0.5) / M)
    yc = np.float32((yM + 0.5) / M)
    X = np.arange(2 * w, dtype=np.float32)[None, :] / np.float32(2)
    Y = np.arange(2 * h, dtype=np.float32)[:, None] / np.float32(2)
    gxi = np.broadcast_to(-(X - xc), (2 * h, 2 * w)).astype(np.float32)
    gyi = np.broadcast_to(-(Y - yc), (2 * h, 2 * w)).astype(np.float32)
    px = np.full((h, w), 7.0, np.float32)
    out, prev = _bic_frame(O, gxi, gyi, px, M, fwhm, 3.0, intW=0, n=2)
    assert out[0, yM, xM] == np.float32(0.5)                    # AVG = CGLH/n, CGLH = 1
    assert out[1, yM, xM] == np.float32(0.5)                    # E[x^2] = CGLH^2/n
    assert prev[yM, xM] == np.float32(1.0)                      # Prev = CGLH/(n-1)
    assert abs(out[3, yM, xM] - 3.5) < 2e-5     # INT = v/n, v = 7 (weights sum to 1 +- ulps)
    out2, _ = _bic_frame(O, -gxi, -gyi, px, M, fwhm, 1.0, intW=0, n=2)
    assert out2[0, yM, xM] == 0.0                               # divergent field
    # a neighbouring pixel sees a field converging elsewhere: strictly between 0 and 1
    assert 0.0 < out[0, yM, xM + 1] < out[0, yM, xM]


def test_rgc_uniform_field_closed_form(O):
    """G = (1, 0) everywhere: inward samples are those with dx < 0 and
    Dk = 1 - |dy|/d, so CGLH = sum_{dx<0} w (1 - |dy|/d) / sum w over the gated window
    (j, i in [-(int)2r, (int)(2r+1)], d != 0, d <= tSO)."""
    M, fwhm, xM, yM = 5, 1.5, 61, 58
    h = w = 26
    gxi = np.ones((2 * h, 2 * w), np.float32)
    gyi = np.zeros((2 * h, 2 * w), np.float32)
    px = np.ones((h, w), np.float32)
    out, _ = _bic_frame(O, gxi, gyi, px, M, fwhm, 1.0, intW=0, n=2)
    k = O.bic_constants(fwhm)
    xc, yc = (xM + 0.5) / M, (yM + 0.5) / M
    num = den = 0.0
    lo, hi = -int(2 * k["radius"]), int(2 * k["radius"] + 1)
    for j in range(lo, hi + 1):
        vy = (int(2 * yc) + j) / 2
        for i in range(lo, hi + 1):
            vx = (int(2 * xc) + i) / 2
            dx, dy = vx - xc, vy - yc
            d = math.hypot(dx, dy)
            if d == 0 or d > k["tSO"]:
                continue
            wgt = (d * math.exp(-d * d / k["tSS"])) ** 4
            den += wgt
            if dx < 0:
                num += wgt * (1 - abs(dy) / d)
    assert abs(out[0, yM, xM] - (num / den) / 2) < 2e-6


def test_rgc_sensitivity_is_a_power(O):
    rng = np.random.default_rng(4)
    h = w = 16
    gxi = rng.normal(size=(2 * h, 2 * w)).astype(np.float32)
    gyi = rng.normal(size=(2 * h, 2 * w)).astype(np.float32)
    px = np.ones((h, w), np.float32)
    o1, _ = _bic_frame(O, gxi, gyi, px, 3, 1.5, 1.0, intW=0, n=2)
    o2, _ = _bic_frame(O, gxi, gyi, px, 3, 1.5, 2.0, intW=0, n=2)
    c1, c2 = 2 * o1[0].astype(np.float64), 2 * o2[0].astype(np.float64)
    assert np.all(c1 >= 0) and np.all(c1 <= 1)
    np.testing.assert_allclose(c2, c1 ** 2, rtol=2e-6, atol=1e-9)


# ---------------------------------------------------------------------------------------
# Temporal cumulants and epilogues (liveSRRF.cl:363-435, 464-505)
# ---------------------------------------------------------------------------------------
def test_constant_in_time_stack_cumulants(O):
    """n copies of one frame: AVG = t, VAR = 0 and the BIC TAC2 quirk
    TAC2 = (n-1) t^2/(n-1)^2 - t^2 = t^2/(n-1) - t^2 (Prev carries 1/(n-1))."""
    from nanoj_esrrf_amd import synth
    f = synth.blinking_stack(1, 20, 18, 30, 3, seed=8)[0]
    n = 6
    st = np.repeat(f[None], n, axis=0).astype(np.float32)
    o = O.OracleLiveSRRF(20, 18, 3, 1.5, 1, n, n, True, False, False)
    o.calculateSRRF(st)
    r = o.readSRRFbuffer().reshape(4, -1).astype(np.float64)
    o1 = O.OracleLiveSRRF(20, 18, 3, 1.5, 1, 1, 1, True, False, False)
    o1.calculateSRRF(st[:1])
    t = o1.out.reshape(4, -1)[0].astype(np.float64)       # one frame, n = 1: t / 1
    np.testing.assert_allclose(r[0], t, rtol=5e-6, atol=1e-6)
    scale = float(np.max(t * t))
    assert np.max(np.abs(r[1])) < 1e-5 * scale
    np.testing.assert_allclose(r[2], t * t / (n - 1) - t * t, rtol=1e-4, atol=1e-5 * scale)


def test_mp_correction_normalises_subposition_means(O):
    from nanoj_esrrf_amd import synth
    st = synth.blinking_stack(5, 16, 14, 30, 3, seed=12).astype(np.float32)
    o = O.OracleLiveSRRF(16, 14, 4, 1.5, 2, 5, 5, True, True, False)
    o.calculateSRRF(st)
    r = o.readSRRFbuffer()
    for c in range(3):
        sub = r[c].reshape(14, 4, 16, 4).mean(axis=(0, 2))
        np.testing.assert_allclose(sub, 1.0, rtol=2e-4)
    assert o.mp.shape == (3 * 16,)


# ---------------------------------------------------------------------------------------
# Fourier (Hartley) upsampling (FourierInterpolation.java)
# ---------------------------------------------------------------------------------------
def test_dht_matches_bruteforce_cas_sum():
    from oracle import fht
    x = np.random.default_rng(5).normal(size=(8, 8))
    np.testing.assert_allclose(fht.dht2(x), fht.dht2_bruteforce(x), atol=1e-9)
    np.testing.assert_allclose(fht.idht2(fht.dht2(x)), x, atol=1e-12)


def test_fht_upsampling_reproduces_bandlimited_samples():
    from oracle import fht
    N, M = 16, 3
    y, x = np.mgrid[0:N, 0:N]
    img = (5 + np.cos(2 * np.pi * 3 * x / N) + 0.5 * np.sin(2 * np.pi * (2 * x + 5 * y) / N))
    up = fht.fht_space_interpolation(img.astype(np.float32), M, mirror=True)
    assert up.shape == (N * M, N * M)
    # normalised inverse: intensity scale 1/M^2; every M-th sample reproduces the input
    np.testing.assert_allclose(up[::M, ::M] * M * M, img, rtol=2e-5, atol=2e-5)


def test_mirror_padding_identity_and_odd_tiles():
    from oracle import fht
    a = np.arange(64, dtype=np.float32).reshape(8, 8)
    assert np.array_equal(fht.mirror_padding_even_square(a), a)
    b = np.arange(30, dtype=np.float32).reshape(5, 6)
    p = fht.mirror_padding_even_square(b)
    assert p.shape == (8, 8)
    # odd tile factor 3: the centre tile is the image, its neighbours are mirror images
    big = np.block([[b[::-1, ::-1], b[::-1], b[::-1, ::-1]], [b[:, ::-1], b, b[:, ::-1]],
                    [b[::-1, ::-1], b[::-1], b[::-1, ::-1]]])
    x0, y0 = (18 - 8) // 2, (15 - 8) // 2
    assert np.array_equal(p, big[y0:y0 + 8, x0:x0 + 8])


# ---------------------------------------------------------------------------------------
# 3-D pieces (live3DSRRF.cl)
# ---------------------------------------------------------------------------------------
def test_3d_alignment_identity_at_zero_calibration(O):
    rng = np.random.default_rng(6)
    w, h, nP, nF = 11, 9, 3, 2
    px = rng.integers(0, 1000, size=(nF, nP, h, w)).astype(np.float32).ravel()
    al = np.empty_like(px)
    O.lib().orc3d_align(px, w, h, nP, nF, np.zeros(3 * nP, np.float32), al)
    assert np.array_equal(al, px)


def test_3d_gradients_exact(O):
    rng = np.random.default_rng(7)
    w, h, nP, nF = 7, 6, 4, 2
    px = rng.integers(0, 1000, size=(nF, nP, h, w)).astype(np.float32)
    gx, gy, gz = (np.empty(px.size, np.float32) for _ in range(3))
    O.lib().orc3d_gradient(px.ravel(), w, h, nP, nF, gx, gy, gz)
    assert np.array_equal(gx.reshape(px.shape), np.diff(px, axis=3, prepend=px[..., :1]))
    assert np.array_equal(gy.reshape(px.shape), np.diff(px, axis=2, prepend=px[:, :, :1]))
    assert np.array_equal(gz.reshape(px.shape), np.diff(px, axis=1, prepend=px[:, :1]))


# ---------------------------------------------------------------------------------------
# Integer bookkeeping (gui/LiveSRRF_optimised_.java) -- exact
# ---------------------------------------------------------------------------------------
def test_rolling_plan_hand_cases(O):
    p = O.rolling_plan(1000, 100, 50, True, 30)
    assert (p["nFrameForSRRF"], p["gap"], p["nSRRFframe"]) == (100, 50, 19)
    # ceil(100/30)=4 loads needed -> ceil(100/4)=25 frames per load, 4 loads of 25
    assert p["nFrameOnGPU"] == 25 and p["nLoads"] == 4
    assert p["plan"][:5] == [(1, 0, 1, 25), (1, 1, 26, 25), (1, 2, 51, 25), (1, 3, 76, 25),
                             (2, 0, 51, 25)]
    q = O.rolling_plan(1000, 0, 0, False, 1000)
    assert (q["nFrameForSRRF"], q["nSRRFframe"], q["nLoads"]) == (1000, 1, 1)
    r = O.rolling_plan(250, 100, 0, True, 33)
    # gap 0 -> no rolling overlap; (int)(150/100)+1 = 2 SR frames; uneven last load
    assert r["gap"] == 100 and r["nSRRFframe"] == 2
    assert [x[3] for x in r["plan"] if x[0] == 1] == [25, 25, 25, 25]
    s = O.rolling_plan(103, 103, 0, False, 10)
    assert s["nFrameOnGPU"] == 10 and [x[3] for x in s["plan"]][-1] == 3


def test_product_rolling_plan_matches_oracle_exhaustively(O, pkg):
    for nS in (1, 7, 100, 101, 513):
        for nU in (0, 1, 5, 50, 100, 600):
            for gap in (0, 1, 3, 25):
                for roll in (False, True):
                    for mx in (1, 3, 16, 1000):
                        a = O.rolling_plan(nS, nU, gap, roll, mx)
                        b = pkg._capi.plan_rolling(nS, nU, gap, roll, mx)
                        assert a == b, (nS, nU, gap, roll, mx)


def test_sweep_plan_order(O):
    s = O.sweep_plan(1.0, 0.5, 5, 1, 1, 4, 100, 150, 3, 500)
    assert s["fwhm"] == [1.0, 1.5, 2.0, 2.5, 3.0] and s["nframes"] == [100, 250, 400]
    assert s["order"][0] == (100, 1, 1.0) and s["order"][4] == (100, 1, 3.0)
    assert s["order"][5] == (100, 2, 1.0) and len(s["order"]) == 60


def test_sweep_axes_match_oracle_plan(O, pkg):
    """Host-side sweep axes (Java float / int arithmetic of ParametersSweep_.grabSettings)."""
    for args in [(1.0, 0.5, 5, 1, 1, 4, 100, 150, 3, 500, False),
                 (0.7, 0.3, 4, 2, 2, 3, 50, 100, 5, 260, True),
                 (1.0, 0.5, 2, 1, 1, 2, 300, 100, 3, 200, False)]:
        f, s, n = pkg.sweep_axes(*args)
        plan = O.sweep_plan(*args[:10], frc=args[10])
        assert f == [float(x) for x in plan["fwhm"]]
        assert s == plan["sensitivity"] and n == plan["nframes"]
    assert pkg.sweep_axes(1.0, 0.5, 5, 1, 1, 4, 100, 150, 3, 500)[2] == [100, 250, 400]
    assert pkg.sweep_axes(1.0, 0.5, 5, 1, 1, 4, 300, 100, 3, 200)[2] == [300][:0]


# ---------------------------------------------------------------------------------------
# 3-D axial magnification (an extension: live3DSRRF.cl has one magnification)
# ---------------------------------------------------------------------------------------
def _oracle_3d(O, planes, M, Mz, mp):
    """planes [n][nP][h][w] as a 3x3 montage through the oracle's 3-D state machine."""
    n, nP, h, w = planes.shape
    mont = np.zeros((n, 3 * h, 3 * w), np.float32)
    for z in range(nP):
        i, j = z % 3, z // 3
        mont[:, j * h:(j + 1) * h, i * w:(i + 1) * w] = planes[:, z]
    zc = np.zeros(nP)
    calib = O.Calib3D(zc, zc, zc, np.arange(nP, dtype=np.float64), np.ones(nP), 0.0)
    o = O.OracleLiveSRRF(3 * w, 3 * h, M, 1.5, 2.0, n, n, True, mp, False, calib3d=calib,
                         axialMagnification=Mz)
    o.calculateSRRF(mont)
    return o.readSRRFbuffer().reshape(4, nP * (Mz or M), h * M, w * M)


@pytest.mark.parametrize("M,Mz", [(2, 1), (3, 2), (2, 4)])
def test_oracle_axial_magnification_plane_centres(O, M, Mz):
    """Output plane 3z+1 at axial magnification 3 Mz has the centre
    (3z + 1.5) / (3 Mz) = (z + 0.5) / Mz of plane z at Mz, bit for bit the same float, and
    every other quantity of a voxel depends on (xM, yM) alone: without MP correction the two
    planes agree exactly.  A z index or centre taken from the lateral magnification fails
    this; Mz = M is the reference itself (the golden 3-D vectors)."""
    rng = np.random.default_rng(M * 10 + Mz)
    planes = rng.integers(100, 900, size=(3, 4, 9, 10)).astype(np.float32)
    a = _oracle_3d(O, planes, M, Mz, False)
    b = _oracle_3d(O, planes, M, 3 * Mz, False)
    nP = planes.shape[1]
    for z in range(nP * Mz):
        assert np.array_equal(b[:, 3 * z + 1], a[:, z]), z
    assert np.array_equal(_oracle_3d(O, planes, M, None, False), _oracle_3d(O, planes, M, M, False))


def test_oracle_axial_mp_map_means_are_one(O):
    """With axial magnification the MP correction divides each (z, y, x) sub-position of the
    M x M x Mz macro-voxel by its mean: afterwards every sub-position averages 1."""
    rng = np.random.default_rng(4)
    planes = rng.integers(100, 900, size=(3, 4, 9, 10)).astype(np.float32)
    M, Mz = 2, 3
    out = _oracle_3d(O, planes, M, Mz, True)
    nP, h, w = 4, 9, 10
    for c in range(3):
        v = out[c].astype(np.float64).reshape(nP, Mz, h, M, w, M)
        means = v.mean(axis=(0, 2, 4))
        assert np.allclose(means, 1.0, atol=1e-4), (c, means)
