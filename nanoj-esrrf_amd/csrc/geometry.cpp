// geometry.cpp -- host-side tap tables of the tiled BIC RGC kernel.
//
// For an unshifted frame the RGC geometry of magnified pixel (xM, yM) depends on the
// sub-pixel phase (xM mod M, yM mod M) and, through the float rounding of the centre
// xc = (float)((xM + 0.5) / M) (liveSRRF.cl:263), on the binade of the macro-pixel index
// X = xM / M: for X >= Xs (Xs a power of two >= 1 - min window offset, so that every
// vx - xc subtraction is exact by Sterbenz), xc = X + round_{ulp(X)}(frac) and every
// quantity the reference derives (dx, dy, distance, gate, weight) is identical for all X of
// one binade.  Below Xs the rounding error of xc is < 2^-21 and one representative is used.
// Each (y-class, x-class, phase) entry lists its taps in the reference's j-then-i order
// (liveSRRF.cl:279-284) with exactly the reference's float arithmetic.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "esrrf_internal.h"
#include "geometry.h"

#pragma STDC FP_CONTRACT OFF

namespace esrrf {

static float cubic_host(float x)
{
    const float a = 0.5f;
    if (x < 0.0f) x = -x;
    float z = 0.0f;
    if (x < 1.0f) {
        volatile float t = x * (-a + 2.0f);
        t = t + (a - 3.0f);
        volatile float zz = x * x;
        zz = zz * t;
        z = zz + 1.0f;
    } else if (x < 2.0f) {
        volatile float t1 = -a * x;
        t1 = t1 * x;
        t1 = t1 * x;
        volatile float t2 = 5.0f * a;
        t2 = t2 * x;
        t2 = t2 * x;
        volatile float t3 = 8.0f * a;
        t3 = t3 * x;
        volatile float zz = t1 + t2;
        zz = zz - t3;
        z = zz + 4.0f * a;
    }
    return z;
}

// distanceWeight = d*exp(-(d*d)/tSS); then w*w*w*w, left to right (liveSRRF.cl:296-297)
static float tap_weight(float d, float tSS)
{
    volatile float e = -(d * d);
    e = e / tSS;
    volatile float w1 = d * std::exp((float)e);
    volatile float W = w1 * w1;
    W = W * w1;
    W = W * w1;
    return W;
}

static int class_of(int X, int xs_log2)
{
    if (X < (1 << xs_log2)) return 0;
    int lg = 0;
    while ((X >> (lg + 1)) > 0) lg++;
    return lg - xs_log2 + 1;
}

static int rep_of(int cls, int xs_log2)
{
    return cls == 0 ? (1 << xs_log2) / 2 : (1 << (xs_log2 + cls - 1));
}

// Range [kmin, kmax] of the 2x-grid offsets kx - 2X (and ky - 2Y) of the taps that pass
// the distance gate d <= tSO, over every class representative and sub-pixel phase, with the
// reference's float arithmetic (the same tests as the table builders below).  The LDS tile
// only has to reach these offsets, which at the Java-text gate lie well inside the window.
static void gated_extent(int M, int R2lo, int R2hi, float tSO, int ncls_x, int ncls_y, int lg,
                         int &kmin, int &kmax)
{
    kmin = 1 << 20;
    kmax = -(1 << 20);
    for (int cy = 0; cy < ncls_y; cy++)
    for (int cx = 0; cx < ncls_x; cx++) {
        const int Xr = rep_of(cx, lg), Yr = rep_of(cy, lg);
        for (int sy = 0; sy < M; sy++) {
            for (int sx = 0; sx < M; sx++) {
                const int xM = Xr * M + sx, yM = Yr * M + sy;
                const float xc = (float)((xM + 0.5) / M);
                const float yc = (float)((yM + 0.5) / M);
                const int bx = (int)(2.0f * xc), by = (int)(2.0f * yc);
                for (int j = R2lo; j <= R2hi; j++) {
                    const float vy = ((float)by + (float)j) / 2.0f;
                    for (int i = R2lo; i <= R2hi; i++) {
                        const float vx = ((float)bx + (float)i) / 2.0f;
                        volatile float dx = vx - xc;
                        volatile float dy = vy - yc;
                        volatile float d2 = dx * dx;
                        volatile float dy2 = dy * dy;
                        d2 = d2 + dy2;
                        const float d = std::sqrt((float)d2);
                        if (!(d != 0.0f && d <= tSO)) continue;
                        const int kx = bx - 2 * Xr + i, ky = by - 2 * Yr + j;
                        kmin = std::min(kmin, std::min(kx, ky));
                        kmax = std::max(kmax, std::max(kx, ky));
                    }
                }
            }
        }
    }
    if (kmin > kmax) kmin = kmax = 0;   // no gated tap at all
}

// LDS tile of an 8x8 macro-pixel block: lanes reach 2x-grid columns 2 lx + k, lx in [0, 8),
// k in [kmin, kmax]; G0 (even) shifts the lowest to >= 0.
static void tile_extent(int kmin, int kmax, int &G0, int &TW)
{
    G0 = std::max(0, 2 * ((-kmin + 1) / 2));
    TW = 15 + kmax + G0;
    if (TW & 1) TW++;
}

bool build_bic_geometry(int w, int h, int M, float radius, float tSS, float tSO, BicGeometry &g)
{
    g.R2lo = -(int)(2.0f * radius);
    g.R2hi = (int)(2.0f * radius + 1.0f);
    if (g.R2hi - g.R2lo > 120) return false;
    int xs = 8;
    while (xs < 1 - g.R2lo) xs *= 2;
    int lg = 0;
    while ((1 << lg) < xs) lg++;
    g.xs_log2 = lg;
    g.ncls_x = class_of(w > 0 ? w - 1 : 0, lg) + 1;
    g.ncls_y = class_of(h > 0 ? h - 1 : 0, lg) + 1;
    int kmin, kmax, TW;
    gated_extent(M, g.R2lo, g.R2hi, tSO, g.ncls_x, g.ncls_y, lg, kmin, kmax);
    tile_extent(kmin, kmax, g.G0, TW);
    g.TW = g.TH = TW;
    g.PW = g.PH = TW / 2;
    g.PP = 24;
    while (g.PP < g.PW) g.PP += 32;
    const int MM = M * M;
    const int planesz = g.PP * g.PH;
    g.taps.clear();
    g.tapk.clear();
    g.tapWd.clear();
    g.phases.assign((size_t)g.ncls_x * g.ncls_y * MM, Phase{});
    g.max_off = 0;
    for (int cy = 0; cy < g.ncls_y; cy++) {
        for (int cx = 0; cx < g.ncls_x; cx++) {
            const int Xr = rep_of(cx, lg), Yr = rep_of(cy, lg);
            const int cls_first = (int)g.taps.size();
            for (int sy = 0; sy < M; sy++) {
                for (int sx = 0; sx < M; sx++) {
                    Phase &P = g.phases[((size_t)cy * g.ncls_x + cx) * MM + sy * M + sx];
                    const int xM = Xr * M + sx, yM = Yr * M + sy;
                    const float xc = (float)((xM + 0.5) / M);
                    const float yc = (float)((yM + 0.5) / M);
                    const int bx = (int)(2.0f * xc), by = (int)(2.0f * yc);
                    P.first_tap = (int)g.taps.size();
                    float wsum = 0.0f, wsumd = 0.0f;
                    for (int j = g.R2lo; j <= g.R2hi; j++) {
                        const float vy = ((float)by + (float)j) / 2.0f;
                        for (int i = g.R2lo; i <= g.R2hi; i++) {
                            const float vx = ((float)bx + (float)i) / 2.0f;
                            volatile float dx = vx - xc;
                            volatile float dy = vy - yc;
                            volatile float d2 = dx * dx;
                            volatile float dy2 = dy * dy;
                            d2 = d2 + dy2;
                            const float d = std::sqrt((float)d2);
                            if (!(d != 0.0f && d <= tSO)) continue;
                            const float W = tap_weight(d, tSS);
                            wsum = wsum + W;
                            const int kxr = bx - 2 * Xr + i;
                            const int kyr = by - 2 * Yr + j;
                            const int gx = kxr + g.G0, gy = kyr + g.G0;
                            Tap t;
                            t.off = ((gy & 1) * 2 + (gx & 1)) * planesz + (gy >> 1) * g.PP + (gx >> 1);
                            t.a = (float)dx / d;
                            t.b = (float)dy / d;
                            t.W = (float)W;
                            if (t.off > g.max_off) g.max_off = t.off;
                            g.taps.push_back(t);
                            g.tapk.push_back((kxr + 128) | ((kyr + 128) << 8));
                            // dW/dtSS of W = d^4 exp(-4 d^2 / tSS)
                            const float Wd = (float)((double)(float)W * 4.0 * (double)d * d /
                                                     ((double)tSS * tSS));
                            g.tapWd.push_back(Wd);
                            wsumd = wsumd + Wd;
                        }
                    }
                    P.ntaps2 = ((int)g.taps.size() - P.first_tap + 1) & ~1;
                    // pad to a multiple of 4 with zero-weight taps (the kernel walks the
                    // list in chunks of 4); they add W * h = 0 and are not part of wsum
                    while (((int)g.taps.size() - P.first_tap) % 4 != 0) {
                        Tap z;
                        z.off = 0;
                        z.a = 0.0f;
                        z.b = 0.0f;
                        z.W = 0.0f;
                        g.taps.push_back(z);
                        g.tapk.push_back(128 | (128 << 8));
                        g.tapWd.push_back(0.0f);
                    }
                    P.ntaps = (int)g.taps.size() - P.first_tap;
                    P.wsumd = wsumd;
                    P.sx = sx;
                    P.sy = sy;
                    P.wsum = wsum;
                    // intensity sample at xM/M - 0.5 (liveSRRF.cl:361)
                    volatile float xi = (float)xM / (float)M;
                    xi = xi - 0.5f;
                    volatile float yi = (float)yM / (float)M;
                    yi = yi - 0.5f;
                    const int u0 = (int)std::floor((float)xi), v0 = (int)std::floor((float)yi);
                    P.du = u0 - Xr;
                    P.dv = v0 - Yr;
                    for (int k = 0; k < 4; k++) {
                        P.wx[k] = cubic_host((float)xi - (float)(u0 - 1 + k));
                        P.wy[k] = cubic_host((float)yi - (float)(v0 - 1 + k));
                    }
                }
            }
            g.max_class_taps = std::max(g.max_class_taps, (int)g.taps.size() - cls_first);
        }
    }
    // spare zero taps so the kernel's prefetch of the next chunks never reads past the table
    for (int k = 0; k < TAP_SPARE; k++) {
        Tap z;
        z.off = 0;
        z.a = z.b = z.W = 0.0f;
        g.taps.push_back(z);
        g.tapk.push_back(128 | (128 << 8));
        g.tapWd.push_back(0.0f);
    }
    // every lane (lx, ly in [0, 8)) must stay inside the 4-plane LDS tile
    if (g.max_off + 7 * g.PP + 7 >= 4 * g.PP * g.PH) return false;
    if (g.PW > g.PP) return false;
    return true;
}

bool build_bic_geometry_multi(int w, int h, int M, const float *radius, const float *tSS,
                              const float *tSO, int nf, BicGeometryMulti &g)
{
    if (nf < 1 || nf > SWEEP_NFG) return false;
    float rmax = radius[0], tmax = tSO[0];
    for (int f = 1; f < nf; f++) {
        rmax = std::max(rmax, radius[f]);
        tmax = std::max(tmax, tSO[f]);
    }
    g.nf = nf;
    g.R2lo = -(int)(2.0f * rmax);
    g.R2hi = (int)(2.0f * rmax + 1.0f);
    if (g.R2hi - g.R2lo > 120) return false;
    int xs = 8;
    while (xs < 1 - g.R2lo) xs *= 2;
    int lg = 0;
    while ((1 << lg) < xs) lg++;
    g.xs_log2 = lg;
    g.ncls_x = class_of(w > 0 ? w - 1 : 0, lg) + 1;
    g.ncls_y = class_of(h > 0 ? h - 1 : 0, lg) + 1;
    int kmin, kmax, TW;
    gated_extent(M, g.R2lo, g.R2hi, tmax, g.ncls_x, g.ncls_y, lg, kmin, kmax);
    tile_extent(kmin, kmax, g.G0, TW);
    g.TW = g.TH = TW;
    g.PW = g.PH = TW / 2;
    g.PP = 24;
    while (g.PP < g.PW) g.PP += 32;
    const int MM = M * M;
    const int planesz = g.PP * g.PH;
    g.taps.clear();
    g.tapk.clear();
    g.phases.assign((size_t)g.ncls_x * g.ncls_y * MM, PhaseM{});
    g.max_off = 0;
    g.max_phase_taps = 0;
    for (int cy = 0; cy < g.ncls_y; cy++) {
        for (int cx = 0; cx < g.ncls_x; cx++) {
            const int Xr = rep_of(cx, lg), Yr = rep_of(cy, lg);
            for (int sy = 0; sy < M; sy++) {
                for (int sx = 0; sx < M; sx++) {
                    PhaseM &P = g.phases[((size_t)cy * g.ncls_x + cx) * MM + sy * M + sx];
                    const int xM = Xr * M + sx, yM = Yr * M + sy;
                    const float xc = (float)((xM + 0.5) / M);
                    const float yc = (float)((yM + 0.5) / M);
                    const int bx = (int)(2.0f * xc), by = (int)(2.0f * yc);
                    P.first_tap = (int)g.taps.size();
                    float wsum[SWEEP_NFG] = {};
                    for (int j = g.R2lo; j <= g.R2hi; j++) {
                        const float vy = ((float)by + (float)j) / 2.0f;
                        for (int i = g.R2lo; i <= g.R2hi; i++) {
                            const float vx = ((float)bx + (float)i) / 2.0f;
                            volatile float dx = vx - xc;
                            volatile float dy = vy - yc;
                            volatile float d2 = dx * dx;
                            volatile float dy2 = dy * dy;
                            d2 = d2 + dy2;
                            const float d = std::sqrt((float)d2);
                            if (!(d != 0.0f && d <= tmax)) continue;
                            TapM t{};
                            for (int f = 0; f < nf; f++) {
                                // the same window and gate as fwhm f's own tables
                                const int lo = -(int)(2.0f * radius[f]);
                                const int hi = (int)(2.0f * radius[f] + 1.0f);
                                if (j < lo || j > hi || i < lo || i > hi || !(d <= tSO[f]))
                                    continue;
                                t.W[f] = tap_weight(d, tSS[f]);
                                wsum[f] = wsum[f] + t.W[f];
                            }
                            const int kxr = bx - 2 * Xr + i;
                            const int kyr = by - 2 * Yr + j;
                            const int gx = kxr + g.G0, gy = kyr + g.G0;
                            t.off = ((gy & 1) * 2 + (gx & 1)) * planesz + (gy >> 1) * g.PP + (gx >> 1);
                            t.a = (float)dx / d;
                            t.b = (float)dy / d;
                            if (t.off > g.max_off) g.max_off = t.off;
                            g.taps.push_back(t);
                            g.tapk.push_back((kxr + 128) | ((kyr + 128) << 8));
                        }
                    }
                    while (((int)g.taps.size() - P.first_tap) % 4 != 0) {
                        g.taps.push_back(TapM{});
                        g.tapk.push_back(128 | (128 << 8));
                    }
                    P.ntaps = (int)g.taps.size() - P.first_tap;
                    g.max_phase_taps = std::max(g.max_phase_taps, P.ntaps);
                    for (int f = 0; f < SWEEP_NFG; f++) P.wsum[f] = f < nf ? wsum[f] : 1.0f;
                    P.sx = sx;
                    P.sy = sy;
                    volatile float xi = (float)xM / (float)M;
                    xi = xi - 0.5f;
                    volatile float yi = (float)yM / (float)M;
                    yi = yi - 0.5f;
                    const int u0 = (int)std::floor((float)xi), v0 = (int)std::floor((float)yi);
                    P.du = u0 - Xr;
                    P.dv = v0 - Yr;
                    for (int k = 0; k < 4; k++) {
                        P.wx[k] = cubic_host((float)xi - (float)(u0 - 1 + k));
                        P.wy[k] = cubic_host((float)yi - (float)(v0 - 1 + k));
                    }
                }
            }
        }
    }
    for (int k = 0; k < TAP_SPARE; k++) {
        g.taps.push_back(TapM{});
        g.tapk.push_back(128 | (128 << 8));
    }
    if (g.max_off + 7 * g.PP + 7 >= 4 * g.PP * g.PH) return false;
    if (g.PW > g.PP) return false;
    return true;
}

// Gated sample count of one whole frame (border gates included), geometry only: the
// algorithmic work unit of the RGC (14 FLOP per sample: G.u 3, G x u 3, |G|^2 3, rsqrt 1,
// |c| r 1, 1 - x 1, weighted accumulate 2).
long long count_active_samples_bic(int w, int h, int M, float radius, float tSO)
{
    const int R2lo = -(int)(2.0f * radius), R2hi = (int)(2.0f * radius + 1.0f);
    long long total = 0;
    const int wM = w * M, hM = h * M;
    std::vector<float> xcs(wM), ycs(hM);
    for (int x = 0; x < wM; x++) xcs[x] = (float)((x + 0.5) / M);
    for (int y = 0; y < hM; y++) ycs[y] = (float)((y + 0.5) / M);
    for (int yM = 0; yM < hM; yM++) {
        const float yc = ycs[yM];
        const int by = (int)(2.0f * yc);
        for (int xM = 0; xM < wM; xM++) {
            const float xc = xcs[xM];
            const int bx = (int)(2.0f * xc);
            for (int j = R2lo; j <= R2hi; j++) {
                const float vy = ((float)by + (float)j) / 2.0f;
                if (!(vy > 0.0f && vy < (float)h)) continue;
                const float dy = vy - yc;
                for (int i = R2lo; i <= R2hi; i++) {
                    const float vx = ((float)bx + (float)i) / 2.0f;
                    if (!(vx > 0.0f && vx < (float)w)) continue;
                    const float dx = vx - xc;
                    const float d = std::sqrt(dx * dx + dy * dy);
                    if (d != 0.0f && d <= tSO) total++;
                }
            }
        }
    }
    return total;
}

}  // namespace esrrf

namespace esrrf {

// ---------------------------------------------------------------------------------------
// Tiled 3-D RGC tables.  For a voxel (xM, yM, zM) without drift the sample points of
// live3DSRRF.cl:366-387 are vx = (floor(2 xc) + i)/2, vy = (floor(2 yc) + j)/2 and
// vz = floor(zc) + k; (dx, dy) depend on the lateral phase and, through the rounding of
// xc = (float)((xM + 0.5)/M), on the binade of X = xM / M, exactly as in the 2-D tables
// (for M a power of two xc is exact and one class serves every X); dz depends on zM alone.
// The LDS tile of a block holds the unit gradients of planes [zlo, zlo + nz) over the 2x-
// grid rectangle its 8x8 macro-pixels reach, each row split into its even and odd columns
// so that the stride-2 sample pattern of a wave reads consecutive float4s.
// ---------------------------------------------------------------------------------------
bool build_3d_geometry(int w, int h, int nP, int M, int Mz, float radius, float tSS,
                       float tSO, Geometry3 &g)
{
    g.R2lo = -(int)(2.0f * radius);
    g.R2hi = (int)(2.0f * radius + 1.0f);
    const int klo = -(int)radius, khi = (int)radius + 1;
    if (M < 1 || nP < 2 || g.R2hi - g.R2lo > 60 || khi - klo > 40) return false;
    int xs = 8;
    while (xs < 1 - g.R2lo) xs *= 2;
    int lg = 0;
    while ((1 << lg) < xs) lg++;
    g.xs_log2 = lg;
    const bool pow2 = (M & (M - 1)) == 0;
    g.ncls_x = pow2 ? 1 : class_of(w > 0 ? w - 1 : 0, lg) + 1;
    g.ncls_y = pow2 ? 1 : class_of(h > 0 ? h - 1 : 0, lg) + 1;
    g.nZM = nP * Mz;
    const int MM = M * M;
    // lane columns 2 lx + kx_rel + G0, lx in [0, 8), kx_rel over the lateral offsets the gated
    // 3-D taps reach (every class, z-slice and phase, the reference's float arithmetic)
    {
        int kmin = 1 << 20, kmax = -(1 << 20);
        for (int cy = 0; cy < g.ncls_y; cy++)
        for (int cx = 0; cx < g.ncls_x; cx++) {
            const int Xr = pow2 ? 8 : rep_of(cx, lg), Yr = pow2 ? 8 : rep_of(cy, lg);
            for (int zM = 0; zM < g.nZM; zM++) {
                const float zc = (float)((zM + 0.5) / Mz);
                const int zi = (int)zc;
                for (int sy = 0; sy < M; sy++)
                for (int sx = 0; sx < M; sx++) {
                    const float xc = (float)((Xr * M + sx + 0.5) / M);
                    const float yc = (float)((Yr * M + sy + 0.5) / M);
                    const int bx = (int)(2.0f * xc), by = (int)(2.0f * yc);
                    for (int j = g.R2lo; j <= g.R2hi; j++) {
                        const float vy = ((float)by + (float)j) / 2.0f;
                        for (int i = g.R2lo; i <= g.R2hi; i++) {
                            const float vx = ((float)bx + (float)i) / 2.0f;
                            for (int k = klo; k <= khi; k++) {
                                const float vz = (float)zi + (float)k;
                                if (!(vz > 0.0f && vz < (float)nP)) continue;
                                volatile float dx = vx - xc;
                                volatile float dy = vy - yc;
                                volatile float dz = vz - zc;
                                volatile float d2 = dx * dx;
                                volatile float t2 = dy * dy;
                                d2 = d2 + t2;
                                t2 = dz * dz;
                                d2 = d2 + t2;
                                const float d = std::sqrt((float)d2);
                                if (!(d != 0.0f && d <= tSO)) continue;
                                const int kx = bx - 2 * Xr + i, ky = by - 2 * Yr + j;
                                kmin = std::min(kmin, std::min(kx, ky));
                                kmax = std::max(kmax, std::max(kx, ky));
                            }
                        }
                    }
                }
            }
        }
        if (kmin > kmax) kmin = kmax = 0;
        int TW;
        tile_extent(kmin, kmax, g.G0, TW);
        g.TW = g.TH = TW;
    }
    g.PW = (g.TW + 1) / 2;
    // row pitch = 4 (mod 8) float4s: rows 2 ly apart land in opposite 128-byte bank halves
    g.P = 2 * g.PW;
    while (g.P % 8 != 4) g.P++;
    g.PZ = g.TH * g.P;

    // planes each z-slice can reach (|dz| <= tSO, 0 < vz < nP)
    std::vector<int> zlo(g.nZM), nz(g.nZM);
    g.NZ = 0;
    for (int zM = 0; zM < g.nZM; zM++) {
        const float zc = (float)((zM + 0.5) / Mz);
        int lo = 1 << 30, hi = -1;
        for (int k = klo; k <= khi; k++) {
            const float vz = (float)((int)zc) + (float)k;
            if (!(vz > 0.0f && vz < (float)nP)) continue;
            if (!(std::fabs(vz - zc) <= tSO + 1e-5f)) continue;   // margin: d may round below |dz|
            lo = std::min(lo, (int)vz);
            hi = std::max(hi, (int)vz);
        }
        zlo[zM] = hi >= lo ? lo : 1;
        nz[zM] = hi >= lo ? hi - lo + 1 : 0;
        g.NZ = std::max(g.NZ, nz[zM]);
    }
    // planes a block of two consecutive z-slices stages: the union of their ranges
    g.NZ2 = 0;
    for (int z2 = 0; z2 < g.nZM; z2 += 2) {
        int lo = 1 << 30, hi = -1;
        for (int zM = z2; zM < std::min(g.nZM, z2 + 2); zM++) {
            if (nz[zM] == 0) continue;
            lo = std::min(lo, zlo[zM]);
            hi = std::max(hi, zlo[zM] + nz[zM] - 1);
        }
        if (hi >= lo) g.NZ2 = std::max(g.NZ2, hi - lo + 1);
    }
    int max_inplane = 0;

    g.tapv.clear();
    g.tapo.clear();
    g.tapk.clear();
    g.phases.assign((size_t)g.ncls_y * g.ncls_x * g.nZM * MM, Phase3{});
    g.max_group_taps = 0;
    int max_off = 0;
    const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int cy = 0; cy < g.ncls_y; cy++) {
        for (int cx = 0; cx < g.ncls_x; cx++) {
            const int Xr = pow2 ? 8 : rep_of(cx, lg), Yr = pow2 ? 8 : rep_of(cy, lg);
            for (int zM = 0; zM < g.nZM; zM++) {
                const float zc = (float)((zM + 0.5) / Mz);
                const int zi = (int)zc;
                Phase3 *PB = &g.phases[(((size_t)cy * g.ncls_x + cx) * g.nZM + zM) * MM];
                for (int sy = 0; sy < M; sy++) {
                    for (int sx = 0; sx < M; sx++) {
                        Phase3 &P = PB[sy * M + sx];
                        const int xM = Xr * M + sx, yM = Yr * M + sy;
                        const float xc = (float)((xM + 0.5) / M);
                        const float yc = (float)((yM + 0.5) / M);
                        const int bx = (int)(2.0f * xc), by = (int)(2.0f * yc);
                        P.first_tap = (int)g.tapv.size();
                        float wsum = 0.0f;
                        for (int j = g.R2lo; j <= g.R2hi; j++) {
                            const float vy = ((float)by + (float)j) / 2.0f;
                            for (int i = g.R2lo; i <= g.R2hi; i++) {
                                const float vx = ((float)bx + (float)i) / 2.0f;
                                for (int k = klo; k <= khi; k++) {
                                    const float vz = (float)zi + (float)k;
                                    if (!(vz > 0.0f && vz < (float)nP)) continue;
                                    volatile float dx = vx - xc;
                                    volatile float dy = vy - yc;
                                    volatile float dz = vz - zc;
                                    volatile float d2 = dx * dx;
                                    volatile float t2 = dy * dy;
                                    d2 = d2 + t2;
                                    t2 = dz * dz;
                                    d2 = d2 + t2;
                                    const float d = std::sqrt((float)d2);
                                    if (!(d != 0.0f && d <= tSO)) continue;
                                    const int kz = (int)vz - zlo[zM];
                                    if (kz < 0 || kz >= nz[zM]) return false;
                                    // distanceWeight (live3DSRRF.cl:383-384), left to right
                                    volatile float e = -(d * d);
                                    e = e / tSS;
                                    volatile float w1 = d * std::exp((float)e);
                                    volatile float W = w1 * w1;
                                    W = W * w1;
                                    W = W * w1;
                                    wsum = wsum + (float)W;
                                    const int kxr = bx - 2 * Xr + i, kyr = by - 2 * Yr + j;
                                    const int gx = kxr + g.G0, gy = kyr + g.G0;
                                    const int off = kz * g.PZ + gy * g.P + (gx & 1) * g.PW + (gx >> 1);
                                    max_off = std::max(max_off, off);
                                    max_inplane = std::max(max_inplane, off - kz * g.PZ);
                                    g.tapv.push_back(make_float4((float)dx / d, (float)dy / d,
                                                                 (float)dz / d, (float)W));
                                    g.tapo.push_back(off);
                                    g.tapk.push_back((kxr + 128) | ((kyr + 128) << 8));
                                }
                            }
                        }
                        while (((int)g.tapv.size() - P.first_tap) % 4 != 0) {
                            g.tapv.push_back(zero4);
                            g.tapo.push_back(0);
                            g.tapk.push_back(128 | (128 << 8));
                        }
                        P.ntaps = (int)g.tapv.size() - P.first_tap;
                        P.wsum = wsum;
                        P.zlo = zlo[zM];
                        P.nz = nz[zM];
                        P.sx = sx;
                        P.sy = sy;
                    }
                }
                for (int g0 = 0; g0 < MM; g0 += 4) {
                    const int gl = std::min(MM, g0 + 4) - 1;
                    g.max_group_taps = std::max(
                        g.max_group_taps, PB[gl].first_tap + PB[gl].ntaps - PB[g0].first_tap);
                }
            }
        }
    }
    for (int k = 0; k < TAP_SPARE; k++) {
        g.tapv.push_back(zero4);
        g.tapo.push_back(0);
        g.tapk.push_back(128 | (128 << 8));
    }
    // every lane's reads (base 2 ly P + lx, lx, ly < 8) stay inside the NZ-plane tile
    if (max_off + 14 * g.P + 7 >= g.NZ * g.PZ) return false;
    // ... and inside the tap's own plane, so a slice may sit at any plane of a larger tile
    g.pair_ok = max_inplane + 14 * g.P + 7 < g.PZ;
    if ((int)(2 * g.PW) > g.P) return false;
    return true;
}

}  // namespace esrrf
