/*
 * esrrf_oracle.c -- CPU restatement of the eSRRF reconstruction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the CPU reference that the HIP engine in
 * nanoj-esrrf_amd/ is compared against.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * It restates, operation for operation and in the same floating-point order,
 * the OpenCL kernels that LiveSRRF_CL drives:
 *   BIC 2D  : resources/liveSRRF.cl
 *   FHT 2D  : resources/liveSRRF_noInterpolation.cl (RobX gradient + RGC in
 *             magnified space; the Fourier upsampling itself is in fht.py)
 *   3D      : resources/live3DSRRF.cl
 * Constants are derived as LiveSRRF_CL.java:304-305,317-320 (BIC/3D) and
 * :491-500 (FHT) do, in float32.
 *
 * Where the OpenCL source writes a double literal (0.5 in `(xM + 0.5) /
 * magnification`, the substituted "0.5" of vxy_offset), this file computes in
 * double exactly as an fp64-capable OpenCL device does.  Built with
 * -ffp-contract=off so every multiply/add rounds separately.
 *
 * Parity pins (see tests/test_oracle_*.py): exact integer gradients, closed
 * forms of the Catmull-Rom/bilinear interpolant (quadratic / linear
 * reproduction), radial-field RGC invariants (CGLH == 1 at the convergence
 * point, 0 for divergent fields), constant-stack cumulants, MP-map mean == 1,
 * and the reference kernels themselves compiled from /root/reference and run
 * through the AMD OpenCL runtime (oracle/_ref, GPU tests).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "esrrf_oracle.h"

/* ------------------------------------------------------------------------- */
/* Java's Float.toString and the OpenCL literal read-back (LiveSRRF_CL.java:318) */
/* ------------------------------------------------------------------------- */

/* value = dig * 10^e */
static float dec_as_float(unsigned long long dig, int e)
{
    char b[48];
    snprintf(b, sizeof b, "%llue%d", dig, e);
    return strtof(b, NULL);
}

static int dec_len(unsigned long long v)
{
    int n = 1;
    while (v >= 10) { v /= 10; n++; }
    return n;
}

/* Float.toString(x) as the Java SE specification states it (java.lang.Double.toString,
 * JDK 19+ wording): among the decimals that round to x, those of minimal length n (and of
 * length 2 as well when n == 1); the one closest to x, even last digit on a tie; written
 * plainly for 1e-3 <= |x| < 1e7, otherwise as d.dddE<n>; always a digit after the point.
 * Assumes the "C" numeric locale (the oracle runs in test processes). */
void orc_java_float_to_string(float x, char *out, int len)
{
    if (isnan(x)) { snprintf(out, len, "NaN"); return; }
    if (isinf(x)) { snprintf(out, len, x > 0 ? "Infinity" : "-Infinity"); return; }
    if (x == 0.0f) { snprintf(out, len, signbit(x) ? "-0.0" : "0.0"); return; }
    const float a = fabsf(x);
    unsigned long long cd[64];
    int ce[64], nc = 0;
    for (int q = 1; q <= 9; q++) {
        char b[64];
        snprintf(b, sizeof b, "%.*e", q - 1, (double)a);
        unsigned long long D = 0;
        const char *c = b;
        for (; *c && *c != 'e'; c++)
            if (*c >= '0' && *c <= '9') D = D * 10 + (unsigned)(*c - '0');
        int e = atoi(c + 1) - (q - 1);
        unsigned long long p10 = 1;
        for (int k = 1; k < q; k++) p10 *= 10;
        unsigned long long trial[4] = {D, D + 1, D - 1, p10 * 10 - 1};
        int te[4] = {e, e, e, e - 1};
        for (int t = 0; t < 4; t++) {
            if (t == 2 && D <= 1) continue;
            if (t == 3 && D != p10) continue;
            unsigned long long v = trial[t];
            int ve = te[t];
            while (v % 10 == 0) { v /= 10; ve++; }
            if (dec_as_float(v, ve) != a) continue;
            cd[nc] = v;
            ce[nc] = ve;
            nc++;
        }
    }
    int n = 100;
    for (int i = 0; i < nc; i++) if (dec_len(cd[i]) < n) n = dec_len(cd[i]);
    int nmax = n == 1 ? 2 : n, bi = -1;
    long double bd = 0;
    for (int i = 0; i < nc; i++) {
        int l = dec_len(cd[i]);
        if (l < n || l > nmax) continue;
        char b[48];
        snprintf(b, sizeof b, "%llue%d", cd[i], ce[i]);
        long double dist = fabsl(strtold(b, NULL) - (long double)a);
        if (bi < 0 || dist < bd || (dist == bd && cd[i] % 2 == 0 && cd[bi] % 2 != 0)) {
            bi = i;
            bd = dist;
        }
    }
    char ds[32];
    snprintf(ds, sizeof ds, "%llu", cd[bi]);
    int nd = (int)strlen(ds), ex = ce[bi] + nd - 1;
    char body[64];
    if (ex >= -3 && ex <= 6) {
        if (ex >= 0) {
            char ip[32], fp[32];
            int k;
            for (k = 0; k <= ex; k++) ip[k] = k < nd ? ds[k] : '0';
            ip[k] = 0;
            if (ex + 1 < nd) snprintf(fp, sizeof fp, "%s", ds + ex + 1);
            else snprintf(fp, sizeof fp, "0");
            snprintf(body, sizeof body, "%s.%s", ip, fp);
        } else {
            snprintf(body, sizeof body, "0.%.*s%s", -ex - 1, "000000", ds);
        }
    } else {
        snprintf(body, sizeof body, "%c.%sE%d", ds[0], nd > 1 ? ds + 1 : "0", ex);
    }
    snprintf(out, len, "%s%s", x < 0 ? "-" : "", body);
}

/* tSO of the BIC and 3-D kernels: `"" + 2 * sigma + 1` (LiveSRRF_CL.java:318) is the
 * text Float.toString(2*sigma) + "1", read by `(float) tSO` (liveSRRF.cl:291) as a double
 * literal cast to float. */
float orc_java_tso(float two_sigma)
{
    char t[80];
    orc_java_float_to_string(two_sigma, t, (int)sizeof t - 2);
    strcat(t, "1");
    return (float)strtod(t, NULL);
}

/* ------------------------------------------------------------------------- */
/* Constants (LiveSRRF_CL.java:304-305,317-320 and :491-500)                  */
/* ------------------------------------------------------------------------- */

void orc_bic_constants(float fwhm, int gradientMag, float *sigma, float *radius,
                       float *tSS, float *tSO)
{
    /* float sigma = fwhm / 2.354f;                         LiveSRRF_CL.java:304 */
    float s = fwhm / 2.354f;
    /* radius = ((float)((int)(gradientMag*2*sigma)))/gradientMag + 1    :305 */
    float gm2s = (float)gradientMag * 2.0f;
    gm2s = gm2s * s;
    float r = ((float)((int)gm2s)) / (float)gradientMag + 1.0f;
    /* "" + 2*sigma*sigma is numeric (multiplication first)            :317 */
    float tss = (2.0f * s) * s;
    /* "" + 2*sigma + 1 concatenates: Float.toString(2*sigma) + "1"     :318 */
    float tso = orc_java_tso(2.0f * s);
    if (sigma) *sigma = s;
    if (radius) *radius = r;
    if (tSS) *tSS = tss;
    if (tSO) *tSO = tso;
}

void orc_fht_constants(float fwhm, int magnification, float *sigma, float *radius,
                       float *tSS, float *tSO)
{
    float M = (float)magnification;
    float s = fwhm / 2.354f;                                  /* :491 */
    float analysisRadius = (float)((int)(2.0f * s)) + 1.0f;  /* :492 */
    /* 2 * sigma*magnification * sigma*magnification, left to right   :497 */
    float tss = 2.0f * s;
    tss = tss * M;
    tss = tss * s;
    tss = tss * M;
    float tso = (2.0f * s + 1.0f) * M;                        /* :498 */
    float r = M * analysisRadius;                             /* :500 */
    if (sigma) *sigma = s;
    if (radius) *radius = r;
    if (tSS) *tSS = tss;
    if (tSO) *tSO = tso;
}

/* ------------------------------------------------------------------------- */
/* Interpolation (liveSRRF.cl:26-35, 39-133)                                 */
/* ------------------------------------------------------------------------- */

float orc_cubic(float x)
{
    const float a = 0.5f;
    if (x < 0.0f) x = -x;
    float z = 0.0f;
    if (x < 1.0f) {
        float t = x * (-a + 2.0f);
        t = t + (a - 3.0f);
        z = x * x;
        z = z * t;
        z = z + 1.0f;
    } else if (x < 2.0f) {
        /* -a*x*x*x + 5*a*x*x - 8*a*x + 4*a, left to right */
        float t1 = -a * x;
        t1 = t1 * x;
        t1 = t1 * x;
        float t2 = 5.0f * a;
        t2 = t2 * x;
        t2 = t2 * x;
        float t3 = 8.0f * a;
        t3 = t3 * x;
        z = t1 + t2;
        z = z - t3;
        z = z + 4.0f * a;
    }
    return z;
}

/* getInterpolatedValue on one w x h plane: Catmull-Rom when the 4x4 support is
 * inside (0 < u0 < w-2 && 0 < v0 < h-2), else bilinear with extrapolation
 * (base clamped to [0, w-2], fraction not clamped).  liveSRRF.cl:39-133. */
float orc_interp(const float *a, int w, int h, float x, float y)
{
    const int u0 = (int)floorf(x);
    const int v0 = (int)floorf(y);
    float q = 0.0f;
    if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2) {
        for (int j = 0; j <= 3; j++) {
            int v = v0 - 1 + j;
            if (v < 0) v = 0;
            if (v > h - 1) v = h - 1;
            float p = 0.0f;
            for (int i = 0; i <= 3; i++) {
                int u = u0 - 1 + i;
                if (u < 0) u = 0;
                if (u > w - 1) u = w - 1;
                float t = a[v * w + u] * orc_cubic(x - (float)u);
                p = p + t;
            }
            float t = p * orc_cubic(y - (float)v);
            q = q + t;
        }
    } else {
        float xb = fminf((float)w - 2.0f, fmaxf(x, 0.0f));
        float yb = fminf((float)h - 2.0f, fmaxf(y, 0.0f));
        int xbase = (int)xb, xbase1 = xbase + 1;
        int ybase = (int)yb, ybase1 = ybase + 1;
        float xf = x - (float)xbase;
        float yf = y - (float)ybase;
        float ll = a[ybase * w + xbase];
        float lr = a[ybase * w + xbase1];
        float ur = a[ybase1 * w + xbase1];
        float ul = a[ybase1 * w + xbase];
        float ua = ul + xf * (ur - ul);
        float la = ll + xf * (lr - ll);
        q = la + yf * (ua - la);
    }
    return q;
}

static inline int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

/* ------------------------------------------------------------------------- */
/* BIC 2D: gradients (liveSRRF.cl:187-233)                                   */
/* ------------------------------------------------------------------------- */

void orc_gradient_2point(const float *px, int w, int h, int nF, float *gx, float *gy)
{
    const long wh = (long)w * h;
    for (int f = 0; f < nF; f++)
        for (int y1 = 0; y1 < h; y1++)
            for (int x1 = 0; x1 < w; x1++) {
                long off = (long)y1 * w + x1 + wh * f;
                int x0 = x1 - 1 < 0 ? 0 : x1 - 1;
                int y0 = y1 - 1 < 0 ? 0 : y1 - 1;
                gx[off] = px[off] - px[(long)y1 * w + x0 + wh * f];
                gy[off] = px[off] - px[(long)y0 * w + x1 + wh * f];
            }
}

void orc_gradient_interp(const float *g, int w, int h, int nF, float *gint)
{
    const int wI = 2 * w, hI = 2 * h;
    const long wh = (long)w * h, whI = (long)wI * hI;
    for (int f = 0; f < nF; f++)
        for (int y = 0; y < hI; y++)
            for (int x = 0; x < wI; x++)
                gint[(long)y * wI + x + f * whI] =
                    orc_interp(g + wh * f, w, h, (float)x / 2.0f, (float)y / 2.0f);
}

/* ------------------------------------------------------------------------- */
/* Temporal accumulation shared by the three kernels                          */
/* (liveSRRF.cl:363-393, liveSRRF_noInterpolation.cl:359-389,                 */
/*  live3DSRRF.cl:459-490)                                                    */
/* prev_scaled=1: BIC quirk, Prev = t/(n-1) (so TAC2 carries 1/(n-1)^2).      */
/* ------------------------------------------------------------------------- */

static inline void accumulate(float *out, long off, long plane, float *prev, float v,
                              float CGLH, int first, int nF, int intWeighting,
                              int prev_scaled)
{
    const float n = (float)nF;
    const float n1 = (float)(nF - 1);
    if (intWeighting == 1) {
        float vc = v * CGLH;
        float e2 = vc * v;
        e2 = e2 * CGLH;
        if (first) {
            out[off] = vc / n;
            out[off + plane] = e2 / n;
            out[off + 2 * plane] = 0.0f;
            out[off + 3 * plane] = v / n;
        } else {
            out[off] = out[off] + vc / n;
            out[off + plane] = out[off + plane] + e2 / n;
            float t = vc * prev[off];
            out[off + 2 * plane] = out[off + 2 * plane] + t / n1;
            out[off + 3 * plane] = out[off + 3 * plane] + v / n;
        }
        prev[off] = prev_scaled ? vc / n1 : vc;
    } else {
        float cc = CGLH * CGLH;
        if (first) {
            out[off] = CGLH / n;
            out[off + plane] = cc / n;
            out[off + 2 * plane] = 0.0f;
            out[off + 3 * plane] = v / n;
        } else {
            out[off] = out[off] + CGLH / n;
            out[off + plane] = out[off + plane] + cc / n;
            float t = CGLH * prev[off];
            out[off + 2 * plane] = out[off + 2 * plane] + t / n1;
            out[off + 3 * plane] = out[off + 3 * plane] + v / n;
        }
        prev[off] = prev_scaled ? CGLH / n1 : CGLH;
    }
}

static inline float apply_sensitivity(float CGLH, float sensitivity)
{
    /* CGLH >= 0 ? pow(CGLH, S) : 0; NaN (0/0) fails the test -> 0 */
    if (CGLH >= 0.0f) return powf(CGLH, sensitivity);
    return 0.0f;
}

/* ------------------------------------------------------------------------- */
/* BIC 2D RGC + intensity + accumulation for one frame (liveSRRF.cl:238-401)  */
/* ------------------------------------------------------------------------- */

void orc_bic_rgc_frame(const orc_bic_cfg *c, const float *px, const float *gxi,
                       const float *gyi, const float *shiftXY, int count0, int count1,
                       float *out, float *prev, int nthreads)
{
    const int w = c->width, h = c->height, M = c->magnification;
    const int wM = w * M, hM = h * M;
    const long whM = (long)wM * hM;
    const int wI = 2 * w, hI = 2 * h;
    const long whI = (long)wI * hI, wh = (long)w * h;
    const float shiftX = shiftXY[count0];
    const float shiftY = shiftXY[count0 + c->nFrameForSRRF];
    const int jlo = -(int)(2.0f * c->radius);
    const int jhi = (int)(2.0f * c->radius + 1.0f);
    const float tSO = c->tSO, tSS = c->tSS;
    const float *GX = gxi + whI * count1;
    const float *GY = gyi + whI * count1;
    const float *P = px + wh * count1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (long offset = 0; offset < whM; offset++) {
        const int yM = (int)(offset / wM);
        const int xM = (int)(offset - (long)yM * wM);
        const float xc = (float)((xM + 0.5) / M + (double)shiftX);
        const float yc = (float)((yM + 0.5) / M + (double)shiftY);
        float CGLH = 0.0f, sumW = 0.0f;
        for (int j = jlo; j <= jhi; j++) {
            float vy = ((float)((int)(2.0f * yc)) + (float)j) / 2.0f;
            if (vy > 0.0f && vy < (float)h) {
                for (int i = jlo; i <= jhi; i++) {
                    float vx = ((float)((int)(2.0f * xc)) + (float)i) / 2.0f;
                    if (vx > 0.0f && vx < (float)w) {
                        float dx = vx - xc, dy = vy - yc;
                        float d2 = dx * dx;
                        d2 = d2 + dy * dy;
                        float d = sqrtf(d2);
                        if (d != 0.0f && d <= tSO) {
                            int gxX = (int)(2.0 * ((double)vx - 0.5) + 1.0);
                            int gxY = (int)(2.0 * ((double)vy - 0.5));
                            int gyX = (int)(2.0 * ((double)vx - 0.5));
                            int gyY = (int)(2.0 * ((double)vy - 0.5) + 1.0);
                            float Gx = GX[clampi(gxX, 0, wI - 1) + clampi(gxY, 0, hI - 1) * wI];
                            float Gy = GY[clampi(gyX, 0, wI - 1) + clampi(gyY, 0, hI - 1) * wI];
                            float dw = d * expf(-(d * d) / tSS);
                            dw = dw * dw * dw * dw;
                            sumW += dw;
                            float gdotr = Gx * dx;
                            gdotr = gdotr + Gy * dy;
                            if (gdotr < 0.0f) {
                                float gm2 = Gx * Gx;
                                gm2 = gm2 + Gy * Gy;
                                float GMag = sqrtf(gm2);
                                float cr = Gy * dx;
                                cr = cr - Gx * dy;
                                float Dk = fabsf(cr) / GMag;
                                if (isnan(Dk)) Dk = d;
                                Dk = 1.0f - Dk / d;
                                CGLH += Dk * dw;
                            }
                        }
                    }
                }
            }
        }
        CGLH /= sumW;
        CGLH = apply_sensitivity(CGLH, c->sensitivity);
        float v = orc_interp(P, w, h, ((float)xM) / (float)M + shiftX - 0.5f,
                             ((float)yM) / (float)M + shiftY - 0.5f);
        accumulate(out, offset, whM, prev, v, CGLH, count0 == 0, c->nFrameForSRRF,
                   c->intWeighting, 1);
    }
}

/* ------------------------------------------------------------------------- */
/* Epilogues (liveSRRF.cl:427-435, 464-505)                                   */
/* ------------------------------------------------------------------------- */

void orc_var(float *out, long n)
{
    for (long o = 0; o < n; o++) {
        const float av2 = out[o] * out[o];
        out[o + n] = out[o + n] - av2;
        out[o + 2 * n] = out[o + 2 * n] - av2;
    }
}

/* 2D MP map: for each of 3 channels and M*M sub-positions, sequential float
 * mean over all w*h macro-pixels (liveSRRF.cl:464-487). */
void orc_mp_map(const float *out, int w, int h, int M, float *mp)
{
    const int wM = w * M;
    const long whM = (long)wM * h * M, wh = (long)w * h;
    for (int o = 0; o < 3 * M * M; o++) {
        int frame = o / (M * M);
        int yMP = (o - frame * M * M) / M;
        int xMP = o - yMP * M - frame * M * M;
        float s = 0.0f;
        for (long i = 0; i < wh; i++) {
            int y = (int)(i / w);
            int x = (int)(i - (long)y * w);
            long off = xMP + (long)x * M + (long)wM * (yMP + (long)y * M) + frame * whM;
            s += out[off];
        }
        mp[o] = s / (float)wh;
    }
}

void orc_mp_correct(float *out, int w, int h, int M, const float *mp)
{
    const int wM = w * M;
    const long whM = (long)wM * h * M;
    for (long o = 0; o < 3 * whM; o++) {
        int frame = (int)(o / whM);
        int yM = (int)((o - frame * whM) / wM);
        int xM = (int)(o - (long)yM * wM - frame * whM);
        int x = xM / M, y = yM / M;
        int omp_ = M * M * frame + M * (yM - y * M) + xM - x * M;
        out[o] = out[o] / mp[omp_];
    }
}

/* ------------------------------------------------------------------------- */
/* FHT mode (liveSRRF_noInterpolation.cl:138-159, 245-393)                     */
/* ------------------------------------------------------------------------- */

void orc_gradient_robx(const float *px, int w, int h, int nF, float *gx, float *gy)
{
    const long wh = (long)w * h;
    for (int f = 0; f < nF; f++) {
        const long fo = wh * f;
        for (int y1 = 0; y1 < h; y1++)
            for (int x1 = 0; x1 < w; x1++) {
                int x0 = x1 - 1 < 0 ? 0 : x1 - 1;
                int y0 = y1 - 1 < 0 ? 0 : y1 - 1;
                long off = (long)y1 * w + x1 + fo;
                float p01 = px[(long)y0 * w + x1 + fo];
                float p10 = px[(long)y1 * w + x0 + fo];
                float p11 = px[(long)y1 * w + x1 + fo];
                float p00 = px[(long)y0 * w + x0 + fo];
                float a = p01 - p10;
                a = a + p11;
                a = a - p00;
                gx[off] = a;
                float b = -p01 + p10;
                b = b + p11;
                b = b - p00;
                gy[off] = b;
            }
    }
}

void orc_fht_rgc_frame(const orc_fht_cfg *c, const float *px, const float *gx,
                       const float *gy, const float *shiftXY, int count0, int count1,
                       float *out, float *prev, int nthreads)
{
    const int W = c->widthM, H = c->heightM;
    const long wh = (long)W * H;
    const float shiftX = shiftXY[count0];
    const float shiftY = shiftXY[count0 + c->nFrameForSRRF];
    const int jlo = -(int)c->radius;
    const int jhi = (int)(c->radius + 1.0f);
    const float *GX = gx + wh * count1, *GY = gy + wh * count1, *P = px + wh * count1;
    const float tSO = c->tSO, tSS = c->tSS;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (long offset = 0; offset < wh; offset++) {
        const int yM = (int)(offset / W);
        const int xM = (int)(offset - (long)yM * W);
        const float xc = (float)((xM + 0.5) + (double)shiftX);
        const float yc = (float)((yM + 0.5) + (double)shiftY);
        float CGLH = 0.0f, sumW = 0.0f;
        for (int j = jlo; j <= jhi; j++) {
            float vy = (float)(((float)((int)((double)yc - 0.5)) + (float)j) + 0.5);
            if (vy > 0.0f && vy < (float)H) {
                for (int i = jlo; i <= jhi; i++) {
                    float vx = (float)(((float)((int)((double)xc - 0.5)) + (float)i) + 0.5);
                    if (vx > 0.0f && vx < (float)W) {
                        float dx = vx - xc, dy = vy - yc;
                        float d2 = dx * dx;
                        d2 = d2 + dy * dy;
                        float d = sqrtf(d2);
                        if (d != 0.0f && d <= tSO) {
                            int gxX = (int)((double)vx - 0.5 + 1.0);
                            int gxY = (int)((double)vy - 0.5);
                            int gyX = (int)((double)vx - 0.5);
                            int gyY = (int)((double)vy - 0.5 + 1.0);
                            float Gx = GX[clampi(gxX, 0, W - 1) + clampi(gxY, 0, H - 1) * (long)W];
                            float Gy = GY[clampi(gyX, 0, W - 1) + clampi(gyY, 0, H - 1) * (long)W];
                            float dw = d * expf(-(d * d) / tSS);
                            dw = dw * dw * dw * dw;
                            sumW += dw;
                            float gdotr = Gx * dx;
                            gdotr = gdotr + Gy * dy;
                            if (gdotr < 0.0f) {
                                float gm2 = Gx * Gx;
                                gm2 = gm2 + Gy * Gy;
                                float GMag = sqrtf(gm2);
                                float cr = Gy * dx;
                                cr = cr - Gx * dy;
                                float Dk = fabsf(cr) / GMag;
                                if (isnan(Dk)) Dk = d;
                                Dk = 1.0f - Dk / d;
                                CGLH += Dk * dw;
                            }
                        }
                    }
                }
            }
        }
        CGLH /= sumW;
        CGLH = apply_sensitivity(CGLH, c->sensitivity);
        float v;
        if (shiftX == 0.0f && shiftY == 0.0f)
            v = P[(long)yM * W + xM];
        else
            v = orc_interp(P, W, H, (float)xM + shiftX, (float)yM + shiftY);
        accumulate(out, offset, wh, prev, v, CGLH, count0 == 0, c->nFrameForSRRF,
                   c->intWeighting, 0);
    }
}

/* ------------------------------------------------------------------------- */
/* 3D mode (live3DSRRF.cl)                                                     */
/* ------------------------------------------------------------------------- */

/* kernelAlignPixels (live3DSRRF.cl:204-224); st = [theta(rad)][shiftX][shiftY] */
void orc3d_align(const float *px, int w, int h, int nP, int nF, const float *st,
                 float *aligned)
{
    const long wh = (long)w * h, whd = wh * nP;
    for (long off = 0; off < whd * nF; off++) {
        int f = (int)(off / whd);
        int z = (int)((off - f * whd) / wh);
        int y = (int)((off - f * whd - z * wh) / w);
        int x = (int)(off - f * whd - z * wh - (long)y * w);
        float cw = (float)(w - 1) / 2.0f, ch = (float)(h - 1) / 2.0f;
        float ct = cosf(st[z]), sn = sinf(st[z]);
        float xr = (float)x - cw, yr = (float)y - ch;
        float xa = xr * ct;
        xa = xa - yr * sn;
        xa = xa + cw;
        xa = xa - st[z + nP];
        float ya = xr * sn;
        ya = ya + yr * ct;
        ya = ya + ch;
        ya = ya - st[z + 2 * nP];
        aligned[off] = orc_interp(px + whd * f + wh * z, w, h, xa, ya);
    }
}

/* calculateGradient_2point 3D (live3DSRRF.cl:229-279) */
void orc3d_gradient(const float *px, int w, int h, int nP, int nF, float *gx, float *gy,
                    float *gz)
{
    const long wh = (long)w * h, whd = wh * nP;
    for (long off = 0; off < whd * nF; off++) {
        int f = (int)(off / whd);
        int z1 = (int)((off - f * whd) / wh);
        int y1 = (int)((off - f * whd - z1 * wh) / w);
        int x1 = (int)(off - f * whd - z1 * wh - (long)y1 * w);
        int x0 = x1 - 1 < 0 ? 0 : x1 - 1;
        int y0 = y1 - 1 < 0 ? 0 : y1 - 1;
        int z0 = z1 - 1 < 0 ? 0 : z1 - 1;
        gx[off] = px[off] - px[x0 + (long)y1 * w + z1 * wh + whd * f];
        gy[off] = px[off] - px[x1 + (long)y0 * w + z1 * wh + whd * f];
        gz[off] = px[off] - px[x1 + (long)y1 * w + z0 * wh + whd * f];
    }
}

/* calculateGradientInterpolation 3D (live3DSRRF.cl:282-320): lateral 2x per plane */
void orc3d_interp(const float *g, int w, int h, int nP, int nF, float *gi)
{
    orc_gradient_interp(g, w, h, nP * nF, gi);
}

static inline float cross_mag(float a1, float a2, float a3, float b1, float b2, float b3)
{
    float G1 = a2 * b3;
    G1 = G1 - a3 * b2;
    float G2 = a3 * b1;
    G2 = G2 - a1 * b3;
    float G3 = a1 * b2;
    G3 = G3 - a2 * b1;
    float s = G1 * G1;
    s = s + G2 * G2;
    s = s + G3 * G3;
    return sqrtf(s);
}

/* calculateRadialGradientConvergence 3D (live3DSRRF.cl:325-498).  The reference has one
 * magnification for x, y and z; with axialMagnification Mz the voxel centre is
 * zc = (zM + 0.5) / Mz over nP * Mz planes (Mz = M is the reference). */
void orc3d_rgc_frame(const orc_3d_cfg *c, const float *px, const float *gxi,
                     const float *gyi, const float *gzi, const float *shiftXY, int count0,
                     int count1, float *out, float *prev, int nthreads)
{
    const int w = c->width, h = c->height, M = c->magnification, nP = c->nPlanes;
    const int Mz = c->axialMagnification > 0 ? c->axialMagnification : M;
    const int wM = w * M, hM = h * M;
    const long whM = (long)wM * hM, whdM = whM * nP * Mz;
    const int wI = 2 * w, hI = 2 * h;
    const long whI = (long)wI * hI;
    const long wh = (long)w * h, whd = wh * nP;
    const float driftX = shiftXY[count0];
    const float driftY = shiftXY[count0 + c->nFrameForSRRF];
    const int jlo = -(int)(2.0f * c->radius);
    const int jhi = (int)(2.0f * c->radius + 1.0f);
    const int klo = -(int)c->radius;
    const int khi = (int)c->radius + 1;
    const float *GX = gxi + whI * nP * count1;
    const float *GY = gyi + whI * nP * count1;
    const float *GZ = gzi + whI * nP * count1;
    const float *P = px + whd * count1;
    const float tSO = c->tSO, tSS = c->tSS;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (long offset = 0; offset < whdM; offset++) {
        const int zM = (int)(offset / whM);
        const int yM = (int)((offset - zM * whM) / wM);
        const int xM = (int)(offset - zM * whM - (long)yM * wM);
        const float xc = (float)((xM + 0.5) / M + (double)driftX);
        const float yc = (float)((yM + 0.5) / M + (double)driftY);
        const float zc = (float)((zM + 0.5) / Mz);
        float CGLH = 0.0f, sumW = 0.0f;
        for (int j = jlo; j <= jhi; j++) {
            float vy = ((float)((int)(2.0f * yc)) + (float)j) / 2.0f;
            if (!(vy > 0.0f && vy < (float)h)) continue;
            for (int i = jlo; i <= jhi; i++) {
                float vx = ((float)((int)(2.0f * xc)) + (float)i) / 2.0f;
                if (!(vx > 0.0f && vx < (float)w)) continue;
                for (int k = klo; k <= khi; k++) {
                    float vz = (float)((int)zc) + (float)k;
                    if (!(vz > 0.0f && vz < (float)nP)) continue;
                    float dx = vx - xc, dy = vy - yc, dz = vz - zc;
                    float d2 = dx * dx;
                    d2 = d2 + dy * dy;
                    d2 = d2 + dz * dz;
                    float d = sqrtf(d2);
                    if (!(d != 0.0f && d <= tSO)) continue;
                    int xg = (int)(2.0 * ((double)vx - 0.5) + 1.0);
                    int yg = (int)(2.0 * ((double)vy - 0.5));
                    int xh = (int)(2.0 * ((double)vx - 0.5));
                    int yh = (int)(2.0 * ((double)vy - 0.5) + 1.0);
                    int zz = clampi((int)vz, 0, nP - 1);
                    float Gx = GX[whI * zz + clampi(yg, 0, hI - 1) * (long)wI + clampi(xg, 0, wI - 1)];
                    float Gy = GY[whI * zz + clampi(yh, 0, hI - 1) * (long)wI + clampi(xh, 0, wI - 1)];
                    float Gz = GZ[whI * zz + clampi(yg, 0, hI - 1) * (long)wI + clampi(xh, 0, wI - 1)];
                    float dw = d * expf(-(d * d) / tSS);
                    dw = dw * dw * dw * dw;
                    sumW += dw;
                    float gdotr = Gx * dx;
                    gdotr = gdotr + Gy * dy;
                    gdotr = gdotr + Gz * dz;
                    if (gdotr < 0.0f) {
                        float gm2 = Gx * Gx;
                        gm2 = gm2 + Gy * Gy;
                        gm2 = gm2 + Gz * Gz;
                        float GMag = sqrtf(gm2);
                        float Dk = cross_mag(dx, dy, dz, Gx, Gy, Gz) / GMag;
                        if (isnan(Dk)) Dk = d;
                        Dk = 1.0f - Dk / d;
                        CGLH += Dk * dw;
                    }
                }
            }
        }
        CGLH /= sumW;
        CGLH = apply_sensitivity(CGLH, c->sensitivity);
        int z0;
        float zcof = zc - 0.5f;
        if (zcof < 0.0f) z0 = 0;
        else if (zcof >= (float)(nP - 1)) z0 = nP - 2;
        else z0 = (int)floorf(zcof);
        float xi = ((float)xM) / (float)M + driftX - 0.5f;
        float yi = ((float)yM) / (float)M + driftY - 0.5f;
        int zm1 = z0 - 1 < 0 ? 0 : z0 - 1;
        int zp2 = z0 + 1 < nP - 1 ? z0 + 1 : nP - 1;
        float v_1 = orc_interp(P + wh * zm1, w, h, xi, yi);
        float v0 = orc_interp(P + wh * z0, w, h, xi, yi);
        float v1 = orc_interp(P + wh * (z0 + 1), w, h, xi, yi);
        float v2 = orc_interp(P + wh * zp2, w, h, xi, yi);
        float z01 = zc - ((float)z0 + 0.5f);
        float a = -0.5f * v_1;
        a = a + 1.5f * v0;
        a = a - 1.5f * v1;
        a = a + 0.5f * v2;
        float b = v_1 - 2.5f * v0;
        b = b + 2.0f * v1;
        b = b - 0.5f * v2;
        float cc = -0.5f * v_1;
        cc = cc + 0.5f * v1;
        float t1 = a * z01;
        t1 = t1 * z01;
        t1 = t1 * z01;
        float t2 = b * z01;
        t2 = t2 * z01;
        float v = t1 + t2;
        v = v + cc * z01;
        v = v + v0;
        accumulate(out, offset, whdM, prev, v, CGLH, count0 == 0, c->nFrameForSRRF,
                   c->intWeighting, 0);
    }
}

/* kernelCalculateMPmap 3D (live3DSRRF.cl:561-588); Mz sub-positions along z */
void orc3d_mp_map(const float *out, int w, int h, int nP, int M, int Mz, float *mp)
{
    const int MM = M * M, MMM = MM * Mz;
    const int wM = w * M;
    const long whM = (long)wM * h * M, whdM = whM * nP * Mz;
    const long wh = (long)w * h, whd = wh * nP;
    for (int o = 0; o < 3 * MMM; o++) {
        int frame = o / MMM;
        int zMP = (o - frame * MMM) / MM;
        int yMP = (o - frame * MMM - zMP * MM) / M;
        int xMP = o - frame * MMM - zMP * MM - yMP * M;
        float s = 0.0f;
        for (long i = 0; i < whd; i++) {
            int z = (int)(i / wh);
            int y = (int)((i - z * wh) / w);
            int x = (int)(i - (long)y * w - z * wh);
            long off = xMP + (long)x * M + (long)wM * (yMP + (long)y * M) +
                       whM * (zMP + (long)z * Mz) + frame * whdM;
            s += out[off];
        }
        mp[o] = s / (float)whd;
    }
}

/* kernelCorrectMPmap 3D (live3DSRRF.cl:591-611) */
void orc3d_mp_correct(float *out, int w, int h, int nP, int M, int Mz, const float *mp)
{
    const int wM = w * M;
    const long whM = (long)wM * h * M, whdM = whM * nP * Mz;
    const int MM = M * M;
    for (long o = 0; o < 3 * whdM; o++) {
        int frame = (int)(o / whdM);
        int zM = (int)((o - frame * whdM) / whM);
        int yM = (int)((o - frame * whdM - zM * whM) / wM);
        int xM = (int)(o - frame * whdM - zM * whM - (long)yM * wM);
        int x = xM / M, y = yM / M, z = zM / Mz;
        int om = MM * Mz * frame + MM * (zM - z * Mz) + M * (yM - y * M) + xM - x * M;
        out[o] = out[o] / mp[om];
    }
}

/* 3x3-montage reshape (LiveSRRF_CL.java:963-994) for one frame.
 * roi[z] selects the montage cell; (p - bg) / intCoeff[z]. */
void orc3d_reshape(const float *frame, int W, int H, int nP, const double *roi,
                   const double *intCoeff, float bg, float *planes)
{
    const int wS = W / 3, hS = H / 3;
    for (int z = 0; z < nP; z++) {
        int i = (int)roi[z] % 3;
        int j = ((int)roi[z] - i) / 3;
        int x = (int)floorf((float)W / 3.0f * (float)i + 0.5f);
        int y = (int)floorf((float)H / 3.0f * (float)j + 0.5f);
        for (int yy = 0; yy < hS; yy++)
            for (int xx = 0; xx < wS; xx++) {
                float p = frame[(long)(y + yy) * W + (x + xx)];
                planes[(long)z * wS * hS + (long)yy * wS + xx] =
                    (p - bg) / (float)intCoeff[z];
            }
    }
}
