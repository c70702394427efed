// k_bic.hip -- 2-D eSRRF with 2x Catmull-Rom gradient interpolation ("BIC" mode).
//
// Semantics follow resources/liveSRRF.cl; the data flow is MI355X-first:
//
//  k_bic_prep   one pass per frame: 2-point gradient (liveSRRF.cl:187-214), 2x Catmull-Rom
//               / bilinear-extrapolated interpolation (liveSRRF.cl:39-133, 217-233), and the
//               two staggered reads the RGC makes at every sample point (liveSRRF.cl:293-294)
//               folded into ONE unit vector per sample point of the 2x grid:
//                   T(n,m) = (GxInt[m-1][n], GyInt[m][n-1]),  U = T/|T|  (0 outside the
//               0 < v < width/height gates, 0 for zero/non-finite T).
//  k_bic_rgc    one block per 8x8 macro-pixel tile; one wave per sub-pixel phase, so the
//               tap geometry (unit direction, distance weight, LDS offset) is wave-uniform
//               and read through the scalar cache.  The block stages each frame's U tile
//               in LDS (split into the four parity planes of the 2x grid so the stride-2
//               sample pattern of a wave is bank-conflict free), evaluates
//                   (1 - |sin|) [cos < 0]  =  1 + med3(-1, -|s|, -c - 1)
//               with c = U.(a,b), s = U x (a,b): 6 VALU ops per sample, no transcendental;
//               interpolates the intensity with wave-uniform Catmull-Rom weights; and keeps
//               the four temporal cumulants of every pixel in VGPRs across all frames of the
//               launch (liveSRRF.cl:363-393).
//  k_generic_rgc  per-pixel restatement used when frames carry drift shifts (geometry then
//               changes frame to frame) or M is outside the tiled kernel's range.
#include <cstdlib>

#include "device_cumulants.h"
#include "device_interp.h"
#include "device_tiles.h"
#include "esrrf_internal.h"

namespace esrrf {

// ---------------------------------------------------------------------------------------
// Prep: raw frame -> unit tap gradients on the 2x grid.
// Block = 16x16 sample points; the gradient tile it needs is 12x12 (image columns
// kx0/2-2 .. kx0/2+9) and the raw tile 13x13.
// ---------------------------------------------------------------------------------------
constexpr int PREP_T = 16;
constexpr int PREP_G = 12;
constexpr int PREP_R = 13;

__device__ __forceinline__ void load_grad_tiles(const float *__restrict__ P, int w, int h,
                                                int gx0, int gy0, float *lraw, float *lgx,
                                                float *lgy)
{
    const int tid = threadIdx.y * PREP_T + threadIdx.x;
    // raw tile: image columns gx0-1 .. gx0+11, clamped
    for (int e = tid; e < PREP_R * PREP_R; e += PREP_T * PREP_T) {
        const int r = e / PREP_R, c = e - r * PREP_R;
        const int ix = min(max(gx0 - 1 + c, 0), w - 1);
        const int iy = min(max(gy0 - 1 + r, 0), h - 1);
        lraw[e] = P[(long)iy * w + ix];
    }
    __syncthreads();
    // 2-point gradient (liveSRRF.cl:199-204) at image (gx0 + c, gy0 + r)
    for (int e = tid; e < PREP_G * PREP_G; e += PREP_T * PREP_T) {
        const int r = e / PREP_G, c = e - r * PREP_G;
        const int x1 = gx0 + c, y1 = gy0 + r;
        float gx = 0.0f, gy = 0.0f;
        if (x1 >= 0 && x1 < w && y1 >= 0 && y1 < h) {
            // the raw tile is edge-clamped, so its column c holds image column
            // max(x1-1, 0) and its row r holds row max(y1-1, 0)
            const float p = lraw[(r + 1) * PREP_R + (c + 1)];
            gx = p - lraw[(r + 1) * PREP_R + c];
            gy = p - lraw[r * PREP_R + (c + 1)];
        }
        lgx[e] = gx;
        lgy[e] = gy;
    }
    __syncthreads();
}

// T / |T|, or 0 for a zero or non-finite T.  Where |T|^2 lies well inside the float range
// (all gradients of 16-bit frames) one reciprocal square root normalises it; elsewhere T is
// first scaled by its largest component so that the square neither overflows nor underflows.
__device__ __forceinline__ float2 unit_or_zero(float tx, float ty)
{
    const float t = fmaf(tx, tx, ty * ty);
    if (t >= 1e-30f && t <= 1e30f) {
        const float r = rsqrtf(t);
        return make_float2(tx * r, ty * r);
    }
    const float m = fmaxf(fabsf(tx), fabsf(ty));
    if (!(m > 0.0f) || !(m <= 3.4e38f)) return make_float2(0.0f, 0.0f);
    const float sx = tx / m, sy = ty / m;
    const float r = rsqrtf(sx * sx + sy * sy);
    return make_float2(sx * r, sy * r);
}

// GxInt / GyInt at an integer point (x, y) of the 2x grid: interp(G, x/2, y/2).  The
// fractional parts are 0 or 1/2, so the Catmull-Rom weights are (0, 1, 0, 0) or
// (-1, 9, 9, -1)/16 (cubic(0) = 1, cubic(1) = cubic(2) = 0, cubic(1/2) = 9/16,
// cubic(3/2) = -1/16 exactly); sums run in the reference's order (liveSRRF.cl:49-57).
__device__ __forceinline__ float gint_at(const float *t, int pitch, int ox, int oy, int w, int h,
                                         int x, int y)
{
    const int u0 = x >> 1, v0 = y >> 1;
    if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2) {
        const float c0 = -0.0625f, c1 = 0.5625f;
        const float wx0 = (x & 1) ? c0 : 0.0f, wx1 = (x & 1) ? c1 : 1.0f;
        const float wx2 = (x & 1) ? c1 : 0.0f, wx3 = (x & 1) ? c0 : 0.0f;
        const float wy[4] = {(y & 1) ? c0 : 0.0f, (y & 1) ? c1 : 1.0f, (y & 1) ? c1 : 0.0f,
                             (y & 1) ? c0 : 0.0f};
        const float *r = t + (v0 - 1 - oy) * pitch + (u0 - 1 - ox);
        float q = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            float p = 0.0f;
            p = fmaf(r[0], wx0, p);
            p = fmaf(r[1], wx1, p);
            p = fmaf(r[2], wx2, p);
            p = fmaf(r[3], wx3, p);
            q = fmaf(p, wy[j], q);
            r += pitch;
        }
        return q;
    }
    return interp_tile(t, pitch, ox, oy, w, h, (float)x / 2.0f, (float)y / 2.0f);
}

__global__ void __launch_bounds__(256) k_bic_prep(const float *__restrict__ px,
                                                  float2 *__restrict__ U, int w, int h)
{
    __shared__ float lraw[PREP_R * PREP_R];
    __shared__ float lgx[PREP_G * PREP_G];
    __shared__ float lgy[PREP_G * PREP_G];
    const int f = blockIdx.z;
    const int kx0 = blockIdx.x * PREP_T, ky0 = blockIdx.y * PREP_T;
    const int gx0 = (kx0 >> 1) - 2, gy0 = (ky0 >> 1) - 2;
    const long wh = (long)w * h;
    load_grad_tiles(px + wh * f, w, h, gx0, gy0, lraw, lgx, lgy);
    const int kx = kx0 + threadIdx.x, ky = ky0 + threadIdx.y;
    const int wI = 2 * w, hI = 2 * h;
    if (kx >= wI || ky >= hI) return;
    float2 u = make_float2(0.0f, 0.0f);
    if (kx >= 1 && ky >= 1) {   // gates 0 < vx, 0 < vy of liveSRRF.cl:282,285
        const float tx = gint_at(lgx, PREP_G, gx0, gy0, w, h, kx, ky - 1);
        const float ty = gint_at(lgy, PREP_G, gx0, gy0, w, h, kx - 1, ky);
        u = unit_or_zero(tx, ty);
    }
    U[(long)f * wI * hI + (long)ky * wI + kx] = u;
}

// ---------------------------------------------------------------------------------------
// k_bic_prep2: the same values as k_bic_prep over 32x32 sample points per block, with the
// interpolation's row sums shared: the 4-tap sum along x of a gradient row depends only on
// the row and the sample's x (u0 and parity), so each block forms it once per (row, x) in
// LDS and every sample point combines 4 of them along y, in the order gint_at uses.
// Border points take gint_at's bilinear-extrapolation branch on the same tile.
// ---------------------------------------------------------------------------------------
constexpr int P2_T = 32;   // sample points per side of a block
constexpr int P2_G = 20;   // gradient tile: image columns kx0/2 - 2 .. kx0/2 + 17

__device__ __forceinline__ void cubic_half_weights(int odd, float &w0, float &w1, float &w2,
                                                   float &w3)
{
    const float c0 = -0.0625f, c1 = 0.5625f;
    w0 = odd ? c0 : 0.0f;
    w1 = odd ? c1 : 1.0f;
    w2 = odd ? c1 : 0.0f;
    w3 = odd ? c0 : 0.0f;
}

__global__ void __launch_bounds__(256) k_bic_prep2(const float *__restrict__ px,
                                                   float2 *__restrict__ U, int w, int h)
{
    __shared__ float lgx[P2_G * P2_G], lgy[P2_G * P2_G];
    __shared__ float rsx[P2_G * P2_T], rsy[P2_G * P2_T];   // row sums [tile row][sample x]
    const int f = blockIdx.z;
    const int kx0 = blockIdx.x * P2_T, ky0 = blockIdx.y * P2_T;
    const int gx0 = (kx0 >> 1) - 2, gy0 = (ky0 >> 1) - 2;
    const int tid = threadIdx.x;
    const float *P = px + (long)w * h * f;
    // 2-point gradient (liveSRRF.cl:199-204) at image (gx0 + c, gy0 + r): p minus its left
    // and upper neighbours, edge-clamped, 0 outside the image.  Each thread reads its points'
    // three pixels straight from the frame and issues every load before the first use.
    constexpr int GK = (P2_G * P2_G + 255) / 256;
    float vp[GK], vl[GK], vu[GK];
    bool ok[GK];
#pragma unroll
    for (int k = 0; k < GK; k++) {
        const int e = tid + k * 256;
        const int r = e / P2_G, c = e - r * P2_G;
        const int x1 = gx0 + c, y1 = gy0 + r;
        ok[k] = e < P2_G * P2_G && x1 >= 0 && x1 < w && y1 >= 0 && y1 < h;
        vp[k] = vl[k] = vu[k] = 0.0f;
        if (ok[k]) {
            const float *row = P + (long)y1 * w;
            vp[k] = row[x1];
            vl[k] = row[max(x1 - 1, 0)];
            vu[k] = P[(long)max(y1 - 1, 0) * w + x1];
        }
    }
#pragma unroll
    for (int k = 0; k < GK; k++) {
        const int e = tid + k * 256;
        if (e < P2_G * P2_G) {
            lgx[e] = ok[k] ? vp[k] - vl[k] : 0.0f;
            lgy[e] = ok[k] ? vp[k] - vu[k] : 0.0f;
        }
    }
    __syncthreads();
    // row sums along x: GxInt samples x = kx, GyInt samples x = kx - 1
    for (int e = tid; e < P2_G * P2_T; e += 256) {
        const int r = e / P2_T, q = e - r * P2_T;
        const int kx = kx0 + q;
        float w0, w1, w2, w3;
        {
            const int X = kx, u0 = X >> 1;
            cubic_half_weights(X & 1, w0, w1, w2, w3);
            const float *t = lgx + r * P2_G + (u0 - 1 - gx0);
            float p = 0.0f;
            p = fmaf(t[0], w0, p);
            p = fmaf(t[1], w1, p);
            p = fmaf(t[2], w2, p);
            p = fmaf(t[3], w3, p);
            rsx[e] = p;
        }
        {
            const int X = kx - 1, u0 = X >> 1;
            cubic_half_weights(X & 1, w0, w1, w2, w3);
            const float *t = lgy + r * P2_G + (u0 - 1 - gx0);
            float p = 0.0f;
            p = fmaf(t[0], w0, p);
            p = fmaf(t[1], w1, p);
            p = fmaf(t[2], w2, p);
            p = fmaf(t[3], w3, p);
            rsy[e] = p;
        }
    }
    __syncthreads();
    const int wI = 2 * w, hI = 2 * h;
    float2 *Uf = U + (long)wI * hI * f;
#pragma unroll
    for (int o = 0; o < (P2_T * P2_T) / 256; o++) {
        const int e = tid + o * 256;
        const int q = e & (P2_T - 1), yy = e / P2_T;
        const int kx = kx0 + q, ky = ky0 + yy;
        if (kx >= wI || ky >= hI) continue;
        float2 u = make_float2(0.0f, 0.0f);
        if (kx >= 1 && ky >= 1) {   // gates 0 < vx, 0 < vy of liveSRRF.cl:282,285
            float tx, ty, w0, w1, w2, w3;
            {   // GxInt at (kx, ky - 1)
                const int X = kx, Y = ky - 1, u0 = X >> 1, v0 = Y >> 1;
                if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2) {
                    cubic_half_weights(Y & 1, w0, w1, w2, w3);
                    const float *c = rsx + (v0 - 1 - gy0) * P2_T + q;
                    tx = 0.0f;
                    tx = fmaf(c[0], w0, tx);
                    tx = fmaf(c[P2_T], w1, tx);
                    tx = fmaf(c[2 * P2_T], w2, tx);
                    tx = fmaf(c[3 * P2_T], w3, tx);
                } else {
                    tx = interp_tile(lgx, P2_G, gx0, gy0, w, h, (float)X / 2.0f, (float)Y / 2.0f);
                }
            }
            {   // GyInt at (kx - 1, ky)
                const int X = kx - 1, Y = ky, u0 = X >> 1, v0 = Y >> 1;
                if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2) {
                    cubic_half_weights(Y & 1, w0, w1, w2, w3);
                    const float *c = rsy + (v0 - 1 - gy0) * P2_T + q;
                    ty = 0.0f;
                    ty = fmaf(c[0], w0, ty);
                    ty = fmaf(c[P2_T], w1, ty);
                    ty = fmaf(c[2 * P2_T], w2, ty);
                    ty = fmaf(c[3 * P2_T], w3, ty);
                } else {
                    ty = interp_tile(lgy, P2_G, gx0, gy0, w, h, (float)X / 2.0f, (float)Y / 2.0f);
                }
            }
            u = unit_or_zero(tx, ty);
        }
        Uf[(long)ky * wI + kx] = u;
    }
}

// Debug read-back of GxInt / GyInt of one frame (LiveSRRF_CL.readGradientBuffers).
__global__ void __launch_bounds__(256) k_bic_gint(const float *__restrict__ px,
                                                  float *__restrict__ gxi,
                                                  float *__restrict__ gyi, int w, int h)
{
    __shared__ float lraw[PREP_R * PREP_R];
    __shared__ float lgx[PREP_G * PREP_G];
    __shared__ float lgy[PREP_G * PREP_G];
    const int kx0 = blockIdx.x * PREP_T, ky0 = blockIdx.y * PREP_T;
    const int gx0 = (kx0 >> 1) - 2, gy0 = (ky0 >> 1) - 2;
    load_grad_tiles(px, w, h, gx0, gy0, lraw, lgx, lgy);
    const int x = kx0 + threadIdx.x, y = ky0 + threadIdx.y;
    if (x >= 2 * w || y >= 2 * h) return;
    gxi[(long)y * 2 * w + x] = gint_at(lgx, PREP_G, gx0, gy0, w, h, x, y);
    gyi[(long)y * 2 * w + x] = gint_at(lgy, PREP_G, gx0, gy0, w, h, x, y);
}

// k_bic_prep4: k_bic_prep2 for four frames at once.  Every LDS tile holds the four frames'
// values as float4, so a point's index arithmetic, gates and interpolation weights are
// formed once for the four frames; each frame's values come from the same operations as in
// k_bic_prep2 (blockIdx.z: frames 4z .. 4z+3, those past nF skipped).
__global__ void __launch_bounds__(256) k_bic_prep4(const float *__restrict__ px,
                                                   float2 *__restrict__ U, int w, int h, int nF)
{
    __shared__ float4 lgx[P2_G * P2_G], lgy[P2_G * P2_G];
    __shared__ float4 rsx[P2_G * P2_T], rsy[P2_G * P2_T];   // row sums [tile row][sample x]
    const int f0 = blockIdx.z * 4;
    const int nf = min(4, nF - f0);
    const int kx0 = blockIdx.x * P2_T, ky0 = blockIdx.y * P2_T;
    const int gx0 = (kx0 >> 1) - 2, gy0 = (ky0 >> 1) - 2;
    const int tid = threadIdx.x;
    const long wh = (long)w * h;
    const float *P = px + wh * f0;
    // 2-point gradient (liveSRRF.cl:199-204) at image (gx0 + c, gy0 + r) of each frame
    constexpr int GK = (P2_G * P2_G + 255) / 256;
    float vp[GK][4], vl[GK][4], vu[GK][4];
    bool ok[GK];
#pragma unroll
    for (int k = 0; k < GK; k++) {
        const int e = tid + k * 256;
        const int r = e / P2_G, c = e - r * P2_G;
        const int x1 = gx0 + c, y1 = gy0 + r;
        ok[k] = e < P2_G * P2_G && x1 >= 0 && x1 < w && y1 >= 0 && y1 < h;
#pragma unroll
        for (int j = 0; j < 4; j++) vp[k][j] = vl[k][j] = vu[k][j] = 0.0f;
        if (ok[k]) {
            const long op = (long)y1 * w + x1;
            const long ol = (long)y1 * w + max(x1 - 1, 0);
            const long ou = (long)max(y1 - 1, 0) * w + x1;
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (j < nf) {
                    vp[k][j] = P[op + j * wh];
                    vl[k][j] = P[ol + j * wh];
                    vu[k][j] = P[ou + j * wh];
                }
        }
    }
#pragma unroll
    for (int k = 0; k < GK; k++) {
        const int e = tid + k * 256;
        if (e < P2_G * P2_G) {
            lgx[e] = ok[k] ? make_float4(vp[k][0] - vl[k][0], vp[k][1] - vl[k][1],
                                         vp[k][2] - vl[k][2], vp[k][3] - vl[k][3])
                           : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            lgy[e] = ok[k] ? make_float4(vp[k][0] - vu[k][0], vp[k][1] - vu[k][1],
                                         vp[k][2] - vu[k][2], vp[k][3] - vu[k][3])
                           : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    __syncthreads();
    // row sums along x: GxInt samples x = kx, GyInt samples x = kx - 1
    auto rowsum = [](const float4 *t, float w0, float w1, float w2, float w3) {
        float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        p.x = fmaf(t[0].x, w0, p.x);
        p.y = fmaf(t[0].y, w0, p.y);
        p.z = fmaf(t[0].z, w0, p.z);
        p.w = fmaf(t[0].w, w0, p.w);
        p.x = fmaf(t[1].x, w1, p.x);
        p.y = fmaf(t[1].y, w1, p.y);
        p.z = fmaf(t[1].z, w1, p.z);
        p.w = fmaf(t[1].w, w1, p.w);
        p.x = fmaf(t[2].x, w2, p.x);
        p.y = fmaf(t[2].y, w2, p.y);
        p.z = fmaf(t[2].z, w2, p.z);
        p.w = fmaf(t[2].w, w2, p.w);
        p.x = fmaf(t[3].x, w3, p.x);
        p.y = fmaf(t[3].y, w3, p.y);
        p.z = fmaf(t[3].z, w3, p.z);
        p.w = fmaf(t[3].w, w3, p.w);
        return p;
    };
    for (int e = tid; e < P2_G * P2_T; e += 256) {
        const int r = e / P2_T, q = e - r * P2_T;
        const int kx = kx0 + q;
        float w0, w1, w2, w3;
        {
            const int X = kx, u0 = X >> 1;
            cubic_half_weights(X & 1, w0, w1, w2, w3);
            rsx[e] = rowsum(lgx + r * P2_G + (u0 - 1 - gx0), w0, w1, w2, w3);
        }
        {
            const int X = kx - 1, u0 = X >> 1;
            cubic_half_weights(X & 1, w0, w1, w2, w3);
            rsy[e] = rowsum(lgy + r * P2_G + (u0 - 1 - gx0), w0, w1, w2, w3);
        }
    }
    __syncthreads();
    const int wI = 2 * w, hI = 2 * h;
    const long whI = (long)wI * hI;
    float2 *Uf = U + whI * f0;
    const float *lgxf = reinterpret_cast<const float *>(lgx);
    const float *lgyf = reinterpret_cast<const float *>(lgy);
#pragma unroll
    for (int o = 0; o < (P2_T * P2_T) / 256; o++) {
        const int e = tid + o * 256;
        const int q = e & (P2_T - 1), yy = e / P2_T;
        const int kx = kx0 + q, ky = ky0 + yy;
        if (kx >= wI || ky >= hI) continue;
        float tx[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ty[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        const bool gated = kx >= 1 && ky >= 1;   // gates 0 < vx, 0 < vy of liveSRRF.cl:282,285
        if (gated) {
            float w0, w1, w2, w3;
            {   // GxInt at (kx, ky - 1)
                const int X = kx, Y = ky - 1, u0 = X >> 1, v0 = Y >> 1;
                if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2) {
                    cubic_half_weights(Y & 1, w0, w1, w2, w3);
                    const float4 *c = rsx + (v0 - 1 - gy0) * P2_T + q;
                    const float4 c0 = c[0], c1 = c[P2_T], c2 = c[2 * P2_T], c3 = c[3 * P2_T];
                    const float a0[4] = {c0.x, c0.y, c0.z, c0.w}, a1[4] = {c1.x, c1.y, c1.z, c1.w};
                    const float a2[4] = {c2.x, c2.y, c2.z, c2.w}, a3[4] = {c3.x, c3.y, c3.z, c3.w};
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        float t = 0.0f;
                        t = fmaf(a0[j], w0, t);
                        t = fmaf(a1[j], w1, t);
                        t = fmaf(a2[j], w2, t);
                        t = fmaf(a3[j], w3, t);
                        tx[j] = t;
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        if (j < nf)
                            tx[j] = interp_tile_s<4>(lgxf + j, P2_G, gx0, gy0, w, h,
                                                     (float)X / 2.0f, (float)Y / 2.0f);
                }
            }
            {   // GyInt at (kx - 1, ky)
                const int X = kx - 1, Y = ky, u0 = X >> 1, v0 = Y >> 1;
                if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2) {
                    cubic_half_weights(Y & 1, w0, w1, w2, w3);
                    const float4 *c = rsy + (v0 - 1 - gy0) * P2_T + q;
                    const float4 c0 = c[0], c1 = c[P2_T], c2 = c[2 * P2_T], c3 = c[3 * P2_T];
                    const float a0[4] = {c0.x, c0.y, c0.z, c0.w}, a1[4] = {c1.x, c1.y, c1.z, c1.w};
                    const float a2[4] = {c2.x, c2.y, c2.z, c2.w}, a3[4] = {c3.x, c3.y, c3.z, c3.w};
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        float t = 0.0f;
                        t = fmaf(a0[j], w0, t);
                        t = fmaf(a1[j], w1, t);
                        t = fmaf(a2[j], w2, t);
                        t = fmaf(a3[j], w3, t);
                        ty[j] = t;
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        if (j < nf)
                            ty[j] = interp_tile_s<4>(lgyf + j, P2_G, gx0, gy0, w, h,
                                                     (float)X / 2.0f, (float)Y / 2.0f);
                }
            }
        }
        const long pix = (long)ky * wI + kx;
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (j < nf)
                Uf[pix + j * whI] = gated ? unit_or_zero(tx[j], ty[j]) : make_float2(0.0f, 0.0f);
    }
}

hipError_t launch_bic_prep(const float *px, float2 *U, int w, int h, int nF, hipStream_t s,
                           int kind)
{
    if (nF <= 0) return hipSuccess;
    if (kind == 1) {   // one 4x4 stencil per sample point (k_bic_prep)
        dim3 grid((2 * w + PREP_T - 1) / PREP_T, (2 * h + PREP_T - 1) / PREP_T, nF);
        hipLaunchKernelGGL(k_bic_prep, grid, dim3(PREP_T, PREP_T), 0, s, px, U, w, h);
        return hipGetLastError();
    }
    if (kind == 2) {   // shared row sums, one frame per block
        dim3 grid((2 * w + P2_T - 1) / P2_T, (2 * h + P2_T - 1) / P2_T, nF);
        hipLaunchKernelGGL(k_bic_prep2, grid, dim3(256), 0, s, px, U, w, h);
        return hipGetLastError();
    }
    dim3 grid((2 * w + P2_T - 1) / P2_T, (2 * h + P2_T - 1) / P2_T, (nF + 3) / 4);
    hipLaunchKernelGGL(k_bic_prep4, grid, dim3(256), 0, s, px, U, w, h, nF);
    return hipGetLastError();
}

hipError_t launch_bic_gint(const float *px, float *gxi, float *gyi, int w, int h, int f,
                           hipStream_t s)
{
    dim3 grid((2 * w + PREP_T - 1) / PREP_T, (2 * h + PREP_T - 1) / PREP_T, 1);
    hipLaunchKernelGGL(k_bic_gint, grid, dim3(PREP_T, PREP_T), 0, s, px + (long)w * h * f,
                       gxi, gyi, w, h);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Tiled RGC kernel
// ---------------------------------------------------------------------------------------
constexpr int RAW_T = 12;   // raw intensity tile: image columns X0-2 .. X0+9
constexpr int RAW_P = 72;   // pitch = 8 (mod 64): the 8x8 lanes of a read hit distinct banks
constexpr int RAW_SZ = RAW_T * RAW_P;
constexpr int XS_PER_SX = RAW_T * 8;   // x row sums of one column phase: tile rows x lx
constexpr int XS_MAX_M = 8;           // frame-pair kernel: M * M <= 16 waves x 4 phases

__device__ __forceinline__ int axis_class(int X0, int xs_log2)
{
    if (X0 < (1 << xs_log2)) return 0;
    return 32 - __clz(X0) - xs_log2;   // floor(log2 X0) - xs_log2 + 1
}


// Intensity of a border pixel (the reference's bilinear-extrapolation regime).
__device__ __forceinline__ float border_intensity(const float *lR, int X0, int Y0,
                                                           int w, int h, int M, int xM, int yM)
{
    const float xf = (float)xM / (float)M - 0.5f;
    const float yf = (float)yM / (float)M - 0.5f;
    return interp_tile(lR, RAW_P, X0 - 2, Y0 - 2, w, h, xf, yf);
}

// One RGC sample: h = med3(-1, -|s|, -c - 1) = (1 - |sin|)[cos < 0] - 1 with c = U.(a,b)
// and s = U x (a,b) (liveSRRF.cl:296-341 with G normalised); acc += W * h.
__device__ __forceinline__ float tap_h(float2 u, const Tap &tp)
{
    float c2 = fmaf(u.x, -tp.a, -1.0f);
    c2 = fmaf(u.y, -tp.b, c2);
    float s = u.y * tp.a;
    s = fmaf(u.x, -tp.b, s);
    return __builtin_amdgcn_fmed3f(-1.0f, -fabsf(s), c2);
}

__device__ __forceinline__ float tap_acc(float2 u, const Tap &tp, float acc)
{
    return fmaf(tp.W, tap_h(u, tp), acc);
}

// RSE adjoint: t = v C^S (C^S without intensity weighting); returns dt/dC, dt/dS.
__device__ __forceinline__ void sens_derivs(float C, float sens, int sens_int, float v,
                                            bool intw, float &t, float &dtdC, float &dtdS)
{
    if (!(C >= 0.0f)) {
        t = dtdC = dtdS = 0.0f;
        return;
    }
    const float Cs = apply_sens(C, sens, sens_int);
    const float vv = intw ? v : 1.0f;
    t = vv * Cs;
    if (C > 0.0f) {
        const float Cs1 = sens_int == 1 ? 1.0f : (sens_int == 2 ? C : powf(C, sens - 1.0f));
        dtdC = vv * sens * Cs1;
        dtdS = t * logf(C);
    } else {
        // C = 0: d(C^S)/dC is 1 for S = 1, 0 for S > 1 and unbounded for S < 1 (the zero
        // subgradient is taken); d(C^S)/dS = 0
        dtdC = sens_int == 1 ? vv : 0.0f;
        dtdS = 0.0f;
    }
}

// PPW: sub-pixel phases per wave; STAGE_U: U-tile elements each thread stages in
// registers for the next frame (0: no staging, tiles are filled after the compute).
template <int PPW, int STAGE_U, bool TLDS, int MODE>
__global__ void __launch_bounds__(1024)
__attribute__((amdgpu_waves_per_eu(PPW <= 5 ? 5 : 4))) k_bic_rgc(const BicRgcArgs a,
                                                 const Tap *__restrict__ taps,
                                                 const Phase *__restrict__ phases,
                                                 const int *__restrict__ tapk)
{
    extern __shared__ __align__(16) unsigned char smem[];
    float2 *lU = reinterpret_cast<float2 *>(smem);
    const int usz = 4 * a.PP * a.PH;
    // U buffers: two when the next frame is staged in registers, one otherwise (then the
    // block refills it between two barriers)
    constexpr int NBUF = STAGE_U > 0 ? 2 : 1;
    constexpr int PH_UNROLL = (PPW <= 4 && MODE != 2) ? PPW : 1;
    float *lR = reinterpret_cast<float *>(lU + NBUF * usz);

    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lx = lane & 7, ly = lane >> 3;
    const int lin = xcd_tile_1d(a.tiles_x * a.tiles_y);
    const int tX = lin % a.tiles_x, tY = lin / a.tiles_x;
    const int X0 = tX << 3, Y0 = tY << 3;
    const int X = X0 + lx, Y = Y0 + ly;
    const int M = a.M, MM = M * M;
    const int w = a.w, h = a.h;
    const Phase *PHS = phases + (axis_class(Y0, a.xs_log2) * a.ncls_x +
                                   axis_class(X0, a.xs_log2)) * MM;
    const bool interior = (2 * X0 + a.R2lo >= 1) && (2 * (X0 + 7) + 1 + a.R2hi <= 2 * w - 1) &&
                          (2 * Y0 + a.R2lo >= 1) && (2 * (Y0 + 7) + 1 + a.R2hi <= 2 * h - 1);
    const bool inimg = (X < w) && (Y < h);
    const long whM = (long)a.wM * a.hM;

    // per-phase state of this lane: AVG, E[x^2], TAC2, INT, Prev, 1/sum(w), bicubic base
    // (separate small arrays: each stays in VGPRs under uniform dynamic indexing)
    // (MODE 2 reuses them as AVG, dAVG/dtSS, dAVG/dS, INT)
    float s_avg[PPW], s_e2[PPW], s_tac[PPW], s_int[PPW], s_prv[PPW];
    float invw[PPW];
    float rwd[PPW];   // MODE 2: sum of gated dW/dtSS over sum of gated W
    int ibase[PPW];   // LDS index of the lane's bicubic support, -1: bilinear border path
#pragma unroll
    for (int p = 0; p < PPW; p++) {
        const int ph = wave + p * a.nw;
        invw[p] = 0.0f;
        rwd[p] = 0.0f;
        ibase[p] = -1;
        s_avg[p] = s_e2[p] = s_tac[p] = s_int[p] = s_prv[p] = 0.0f;
        if (ph < MM) {
            const Phase P = PHS[ph];
            float ws = P.wsum, wsd = P.wsumd;
            if (!interior) {
                // border: sum only the taps inside the 0 < v < size gates, reference order
                ws = 0.0f;
                wsd = 0.0f;
                for (int t = 0; t < P.ntaps; t++) {
                    const int k = tapk[P.first_tap + t];
                    const int kx = 2 * X + (k & 255) - 128;
                    const int ky = 2 * Y + ((k >> 8) & 255) - 128;
                    if (kx >= 1 && kx <= 2 * w - 1 && ky >= 1 && ky <= 2 * h - 1) {
                        ws += taps[P.first_tap + t].W;
                        if (MODE == 2) wsd += a.tapWd[P.first_tap + t];
                    }
                }
            }
            invw[p] = 1.0f / ws;
            if (MODE == 2) rwd[p] = wsd * invw[p];
            const int u0 = X + P.du, v0 = Y + P.dv;
            if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2)
                ibase[p] = (v0 - 1 - (Y0 - 2)) * RAW_P + (u0 - 1 - (X0 - 2));
            if (MODE != 1 && !a.first_init && inimg) {
                const long pix = (long)(Y * M + P.sy) * a.wM + (X * M + P.sx);
                s_avg[p] = a.out[pix];
                s_e2[p] = a.out[pix + whM];
                s_tac[p] = a.out[pix + 2 * whM];
                s_int[p] = a.out[pix + 3 * whM];
                s_prv[p] = a.prev[pix];
            }
        }
    }

    const int wI = 2 * w, hI = 2 * h;
    const long whI = (long)wI * hI, wh = (long)w * h;
    const int kxo = 2 * X0 - a.G0, kyo = 2 * Y0 - a.G0;
    const int planesz = a.PP * a.PH;
    const int base = ly * a.PP + lx;
    const float rn = a.rn, rn1 = a.rn1;
    const bool intw = a.int_weighting != 0;
    const bool prev_scaled = a.tac_prev_scaled != 0;

    // Per-thread staging plan of the per-frame tiles, computed once: U-tile element
    // k*nthr + tid is taken plane-major (consecutive threads write consecutive LDS words),
    // goff < 0 marks a zero (outside the 2x grid) or unused slot.
    const int pw = a.TW >> 1;
    const int nU = 4 * a.PH * pw;
    constexpr int SU = STAGE_U > 0 ? STAGE_U : 1;
    int goff[SU], loff[SU];
#pragma unroll
    for (int k = 0; k < SU; k++) {
        const int e = tid + k * nthr;
        goff[k] = -2;
        loff[k] = 0;
        if (e < nU) {
            const int pl = e / (a.PH * pw);
            const int r = (e - pl * a.PH * pw) / pw;
            const int c = e - pl * a.PH * pw - r * pw;
            const int gy = 2 * r + (pl >> 1), gx = 2 * c + (pl & 1);
            const int kx = kxo + gx, ky = kyo + gy;
            loff[k] = pl * planesz + r * a.PP + c;
            goff[k] = (gy < a.TH && gx < a.TW && kx >= 0 && kx < wI && ky >= 0 && ky < hI)
                          ? ky * wI + kx : -1;
        }
    }
    const bool staged = STAGE_U > 0 && nthr * STAGE_U >= nU && nthr >= RAW_T * RAW_T;
    int roff = -1, rl = 0;
    if (tid < RAW_T * RAW_T) {
        const int r = tid / RAW_T, c = tid - (tid / RAW_T) * RAW_T;
        roff = min(max(Y0 - 2 + r, 0), h - 1) * w + min(max(X0 - 2 + c, 0), w - 1);
        rl = r * RAW_P + c;
    }


    // TLDS: this block's tap tables (every phase of its x/y class, contiguous on the host)
    // are staged once in LDS and read with uniform-address ds_read_b128 in the tap loop,
    // so the loop's loads are all in-order LDS traffic
    Tap *lT = reinterpret_cast<Tap *>(lR + NBUF * RAW_SZ);
    float *lWd = reinterpret_cast<float *>(lT + a.max_class_taps + TAP_SPARE);   // MODE 2
    // per-element plan of the U-tile fill (computed once per block)
    int *lPlan = TLDS ? reinterpret_cast<int *>(lWd + (MODE == 2 ? a.max_class_taps + TAP_SPARE
                                                                 : 0))
                      : reinterpret_cast<int *>(lT);
    // direct (unstaged) tile fill: used for the first frame and when a block has too few
    // threads to stage a whole tile in registers
    auto fill_direct = [&](int f, float2 *dU, float *dR) {
        const float2 *Uf = a.U + whI * f;
        for (int e = tid; e < nU; e += nthr) {
            // packed plan: LDS index (15 bits), 2x-grid row (8) and column (8) in the tile
            const int pk = lPlan[e];
            const int gy = (pk >> 15) & 255, gx = (pk >> 23) & 255;
            const int kx = kxo + gx, ky = kyo + gy;
            float2 v = make_float2(0.0f, 0.0f);
            if (gy < a.TH && gx < a.TW && kx >= 0 && kx < wI && ky >= 0 && ky < hI)
                v = Uf[(long)ky * wI + kx];
            dU[pk & 0x7fff] = v;
        }
        const float *Pf = a.px + wh * f;
        if (roff >= 0) dR[rl] = Pf[roff];
        for (int e = tid + nthr; e < RAW_T * RAW_T; e += nthr) {
            const int r = e / RAW_T, c = e - r * RAW_T;
            dR[r * RAW_P + c] = Pf[(long)min(max(Y0 - 2 + r, 0), h - 1) * w +
                                   min(max(X0 - 2 + c, 0), w - 1)];
        }
    };

    const int tcls0 = PHS[0].first_tap;
    if (TLDS) {
        const int tcls1 = PHS[MM - 1].first_tap + PHS[MM - 1].ntaps;
        for (int e = tid; e < tcls1 - tcls0 + TAP_SPARE; e += nthr) lT[e] = taps[tcls0 + e];
        if (MODE == 2)
            for (int e = tid; e < tcls1 - tcls0 + TAP_SPARE; e += nthr)
                lWd[e] = a.tapWd[tcls0 + e];
    }
    for (int e = tid; e < nU; e += nthr) {
        const int pl = e / (a.PH * pw);
        const int r = (e - pl * a.PH * pw) / pw;
        const int c = e - pl * a.PH * pw - r * pw;
        const int gy = 2 * r + (pl >> 1), gx = 2 * c + (pl & 1);
        lPlan[e] = (pl * planesz + r * a.PP + c) | (gy << 15) | (gx << 23);
    }
    __syncthreads();
    fill_direct(a.frame0, lU, lR);
    __syncthreads();

    for (int fi = 0; fi < a.nframes; fi++) {
        const int cur = NBUF == 2 ? (fi & 1) : 0;
        const bool more = fi + 1 < a.nframes;
        // issue the global loads of the next frame's tiles; they land during the compute
        float2 su[SU];
        float sr = 0.0f;
        if (more && staged) {
            const float2 *Un = a.U + whI * (a.frame0 + fi + 1);
#pragma unroll
            for (int k = 0; k < SU; k++)
                su[k] = goff[k] >= 0 ? Un[goff[k]] : make_float2(0.0f, 0.0f);
            if (roff >= 0) sr = a.px[wh * (a.frame0 + fi + 1) + roff];
        }
        // (buffer selection by offset arithmetic on the __shared__ base keeps the LDS
        // address space visible to the compiler: ds_read, not flat loads)
        const float2 *cU = lU + cur * usz;
        const float *cR = lR + cur * RAW_SZ;
        const bool first = (fi == 0) && a.first_init;
        // (a.zero is 0 at run time; the compiler cannot prove it, so the per-phase table
        // reads below stay inside the frame loop instead of being hoisted PPW times)
        const int zf = a.zero * fi;
        const Phase *PHf = PHS + zf;
        const Tap *TAPf = taps + zf;

        // Unrolled, the per-phase state (s_avg[p] ...) has static register indices; rolled,
        // every access becomes a chain of selects. Unroll where the registers fit.
#pragma unroll PH_UNROLL
        for (int p = 0; p < PPW; p++) {
            const int ph = wave + p * a.nw;
            if (ph < MM) {
                const Phase P = PHf[ph];
                // Tap lists are padded to a multiple of 4 with W = 0 taps and the table
                // ends with one spare chunk, so the constants of chunk c+1 can always be
                // fetched (scalar cache) while the LDS reads of chunk c are in flight.
                float acc = 0.0f, acc2 = 0.0f;
                if (MODE == 2) {
                    // adjoint: acc2 = sum dW/dtSS * h beside acc (taps and dW/dtSS from LDS
                    // when staged there, else from the global tables)
                    const Tap *Tl = TLDS ? lT + (P.first_tap - tcls0) : TAPf + P.first_tap;
                    const float *Wl = TLDS ? lWd + (P.first_tap - tcls0) : a.tapWd + P.first_tap;
                    for (int t = 0; t < P.ntaps; t += 4) {
                        const Tap c0 = Tl[t], c1 = Tl[t + 1], c2 = Tl[t + 2], c3 = Tl[t + 3];
                        const float4 wd = *reinterpret_cast<const float4 *>(Wl + t);
                        const float h0 = tap_h(cU[base + c0.off], c0);
                        const float h1 = tap_h(cU[base + c1.off], c1);
                        const float h2 = tap_h(cU[base + c2.off], c2);
                        const float h3 = tap_h(cU[base + c3.off], c3);
                        acc = fmaf(c0.W, h0, acc);
                        acc2 = fmaf(wd.x, h0, acc2);
                        acc = fmaf(c1.W, h1, acc);
                        acc2 = fmaf(wd.y, h1, acc2);
                        acc = fmaf(c2.W, h2, acc);
                        acc2 = fmaf(wd.z, h2, acc2);
                        acc = fmaf(c3.W, h3, acc);
                        acc2 = fmaf(wd.w, h3, acc2);
                    }
                } else if (TLDS) {
                    const Tap *Tl = lT + (P.first_tap - tcls0);
                    for (int t = 0; t < P.ntaps; t += 4) {
                        const Tap c0 = Tl[t], c1 = Tl[t + 1], c2 = Tl[t + 2], c3 = Tl[t + 3];
                        const float2 u0 = cU[base + c0.off];
                        const float2 u1 = cU[base + c1.off];
                        const float2 u2 = cU[base + c2.off];
                        const float2 u3 = cU[base + c3.off];
                        acc = tap_acc(u0, c0, acc);
                        acc = tap_acc(u1, c1, acc);
                        acc = tap_acc(u2, c2, acc);
                        acc = tap_acc(u3, c3, acc);
                    }
                } else {
                    const Tap *T = TAPf + P.first_tap;
                    Tap c0 = T[0], c1 = T[1], c2 = T[2], c3 = T[3];
                    const int nch = P.ntaps >> 2;
                    for (int c = 0; c < nch; c++) {
                        const float2 u0 = cU[base + c0.off];
                        const float2 u1 = cU[base + c1.off];
                        const float2 u2 = cU[base + c2.off];
                        const float2 u3 = cU[base + c3.off];
                        T += 4;
                        const Tap n0 = T[0], n1 = T[1], n2 = T[2], n3 = T[3];
                        acc = tap_acc(u0, c0, acc);
                        acc = tap_acc(u1, c1, acc);
                        acc = tap_acc(u2, c2, acc);
                        acc = tap_acc(u3, c3, acc);
                        c0 = n0;
                        c1 = n1;
                        c2 = n2;
                        c3 = n3;
                    }
                }
                const float C0 = (P.wsum + acc) * invw[p];
                const float C = MODE == 0 ? apply_sens(C0, a.sens, a.sens_int) : C0;

                // interpolated intensity at (xM/M - 0.5, yM/M - 0.5) (liveSRRF.cl:361)
                float v;
                if (ibase[p] >= 0) {
                    const float *row = cR + ibase[p];
                    float q = 0.0f;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        float pp = 0.0f;
                        pp = fmaf(row[j * RAW_P + 0], P.wx[0], pp);
                        pp = fmaf(row[j * RAW_P + 1], P.wx[1], pp);
                        pp = fmaf(row[j * RAW_P + 2], P.wx[2], pp);
                        pp = fmaf(row[j * RAW_P + 3], P.wx[3], pp);
                        q = fmaf(pp, P.wy[j], q);
                    }
                    v = q;
                } else {
                    v = border_intensity(cR, X0, Y0, w, h, M, (X + zf) * M + P.sx,
                                         (Y + zf) * M + P.sy);
                }

                if (MODE == 1) {
                    if (inimg) {
                        const long e = (((long)(a.frame0 + fi) * MM + ph) * h + Y) * w + X;
                        a.emitC[e] = C0;
                        a.emitV[e] = v;
                    }
                    continue;
                }
                if (MODE == 2) {
                    // dC/dtSS of C = (wsum + acc) / sum(W) with sum(W) over the gated taps
                    const float Cd = (P.wsumd + acc2) * invw[p] - C0 * rwd[p];
                    float t, dtdC, dtdS;
                    sens_derivs(C0, a.sens, a.sens_int, v, intw, t, dtdC, dtdS);
                    // a pixel without gated samples (sum(W) = 0) has C0 = 0/0 and the
                    // constant output 0 (liveSRRF.cl:358-360): no derivative, rather than
                    // 0 x NaN
                    const float dtdtss = dtdC != 0.0f ? dtdC * Cd : 0.0f;
                    if (first) {
                        s_avg[p] = t * rn;
                        s_e2[p] = dtdtss * rn;
                        s_tac[p] = dtdS * rn;
                        s_int[p] = v * rn;
                    } else {
                        s_avg[p] = s_avg[p] + t * rn;
                        s_e2[p] = s_e2[p] + dtdtss * rn;
                        s_tac[p] = s_tac[p] + dtdS * rn;
                        s_int[p] = s_int[p] + v * rn;
                    }
                    continue;
                }
                // temporal accumulation (liveSRRF.cl:363-393), every rounding spelled out
                // (the frame-pair kernel's cumulate() does the same operations)
                const float t = intw ? __fmul_rn(v, C) : C;
                const float e2 = intw ? __fmul_rn(__fmul_rn(t, v), C) : __fmul_rn(C, C);
                if (first) {
                    s_avg[p] = __fmul_rn(t, rn);
                    s_e2[p] = __fmul_rn(e2, rn);
                    s_tac[p] = 0.0f;
                    s_int[p] = __fmul_rn(v, rn);
                } else {
                    s_avg[p] = fmaf(t, rn, s_avg[p]);
                    s_e2[p] = fmaf(e2, rn, s_e2[p]);
                    s_tac[p] = fmaf(__fmul_rn(t, s_prv[p]), rn1, s_tac[p]);
                    s_int[p] = fmaf(v, rn, s_int[p]);
                }
                s_prv[p] = prev_scaled ? __fmul_rn(t, rn1) : t;
            }
        }
        if (NBUF == 2) {
            if (more) {
                float2 *nU_ = lU + (cur ^ 1) * usz;
                float *nR_ = lR + (cur ^ 1) * RAW_SZ;
                if (staged) {
#pragma unroll
                    for (int k = 0; k < SU; k++)
                        if (goff[k] != -2) nU_[loff[k]] = su[k];
                    if (roff >= 0) nR_[rl] = sr;
                } else {
                    fill_direct(a.frame0 + fi + 1, nU_, nR_);
                }
            }
            __syncthreads();
        } else if (more) {
            __syncthreads();   // every wave is done with this frame's tiles
            fill_direct(a.frame0 + fi + 1, lU, lR);
            __syncthreads();
        }
    }

    if (MODE != 1 && inimg) {
#pragma unroll
        for (int p = 0; p < PPW; p++) {
            const int ph = wave + p * a.nw;
            if (ph >= MM) continue;
            const Phase P = PHS[ph];
            const long pix = (long)(Y * M + P.sy) * a.wM + (X * M + P.sx);
            a.out[pix] = s_avg[p];
            a.out[pix + whM] = s_e2[p];
            a.out[pix + 2 * whM] = s_tac[p];
            a.out[pix + 3 * whM] = s_int[p];
            if (MODE == 0) a.prev[pix] = s_prv[p];
        }
    }
}

// ---------------------------------------------------------------------------------------
// Frame-pair variant of the tiled RGC (modes 0 and 1, tap tables in LDS).  The block holds
// two consecutive frames: the U tile interleaved as float4 (Ux, Uy of frame f, Ux, Uy of
// frame f+1) and the raw tiles as two planes RAW_SZ apart.  One uniform tap read, one
// address add and one 16-byte sample read then serve both frames, and there is one pair of
// barriers per two frames.  The cumulants are updated frame by frame in order, so results
// are those of k_bic_rgc bit for bit.
// ---------------------------------------------------------------------------------------
template <bool INTW>
__device__ __forceinline__ void cumulate(float &avg, float &e2s, float &tac, float &in,
                                         float &prv, float C0, float v, bool first,
                                         const BicRgcArgs &a)
{
    cumulate_bic<INTW>(avg, e2s, tac, in, prv, C0, v, first, a.sens, a.sens_int, a.rn, a.rn1);
}

// TSM: the tap tables are read with scalar loads from global memory (wave-uniform
// addresses, held in SGPRs) instead of from an LDS copy.  STG > 0: the next frame pair's
// tiles are loaded into registers (STG elements per thread) before the current pair is
// evaluated and only stored to LDS between the barriers.
template <int PPW, int MODE, bool INTW, bool TSM, int STG>
__global__ void __launch_bounds__(1024)
__attribute__((amdgpu_waves_per_eu(4))) k_bic_rgc2(const BicRgcArgs a,
                                                   const Tap *__restrict__ taps,
                                                   const Phase *__restrict__ phases,
                                                   const int *__restrict__ tapk)
{
    extern __shared__ __align__(16) unsigned char smem[];
    float4 *lU = reinterpret_cast<float4 *>(smem);
    const int usz = 4 * a.PP * a.PH;
    float2 *lR = reinterpret_cast<float2 *>(lU + usz);   // raw tiles of f and f+1, interleaved
    float4 *lXW = reinterpret_cast<float4 *>(lR + RAW_SZ);   // [M] Catmull-Rom x weights
    float2 *lXS = reinterpret_cast<float2 *>(lXW + a.M);    // [M][RAW_T][8] x row sums
    int *lXD = reinterpret_cast<int *>(lXS + a.M * XS_PER_SX);   // [M] first tile column - lx
    // (every region above is a multiple of 16 bytes long, so the tap table is 16-byte
    // aligned and read with one 16-byte LDS load per tap)
    Tap *lT = static_cast<Tap *>(__builtin_assume_aligned(lXD + XS_MAX_M, 16));
    int *lPlan = TSM ? reinterpret_cast<int *>(lT)
                     : reinterpret_cast<int *>(lT + a.max_class_taps + TAP_SPARE);
    constexpr int PH_UNROLL = PPW <= 4 ? PPW : 1;

    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lx = lane & 7, ly = lane >> 3;
    const int lin = xcd_tile_1d(a.tiles_x * a.tiles_y);
    const int tX = lin % a.tiles_x, tY = lin / a.tiles_x;
    const int X0 = tX << 3, Y0 = tY << 3;
    const int X = X0 + lx, Y = Y0 + ly;
    const int M = a.M, MM = M * M;
    const int w = a.w, h = a.h;
    const Phase *PHS = phases + (axis_class(Y0, a.xs_log2) * a.ncls_x +
                                   axis_class(X0, a.xs_log2)) * MM;
    const bool interior = (2 * X0 + a.R2lo >= 1) && (2 * (X0 + 7) + 1 + a.R2hi <= 2 * w - 1) &&
                          (2 * Y0 + a.R2lo >= 1) && (2 * (Y0 + 7) + 1 + a.R2hi <= 2 * h - 1);
    const bool inimg = (X < w) && (Y < h);
    const long whM = (long)a.wM * a.hM;

    float s_avg[PPW], s_e2[PPW], s_tac[PPW], s_int[PPW], s_prv[PPW];
    float invw[PPW];
    int xbase[PPW];
#pragma unroll
    for (int p = 0; p < PPW; p++) {
        const int ph = wave + p * a.nw;
        invw[p] = 0.0f;
        xbase[p] = -1;
        s_avg[p] = s_e2[p] = s_tac[p] = s_int[p] = s_prv[p] = 0.0f;
        if (ph < MM) {
            const Phase P = PHS[ph];
            float ws = P.wsum;
            if (!interior) {
                // border: sum only the taps inside the 0 < v < size gates, reference order
                ws = 0.0f;
                for (int t = 0; t < P.ntaps; t++) {
                    const int k = tapk[P.first_tap + t];
                    const int kx = 2 * X + (k & 255) - 128;
                    const int ky = 2 * Y + ((k >> 8) & 255) - 128;
                    if (kx >= 1 && kx <= 2 * w - 1 && ky >= 1 && ky <= 2 * h - 1)
                        ws += taps[P.first_tap + t].W;
                }
            }
            invw[p] = 1.0f / ws;
            // bicubic regime: the four x row sums of the phase's column phase sx, tile rows
            // v0 - 1 .. v0 + 2 (tile row 0 = image row Y0 - 2)
            const int u0 = X + P.du, v0 = Y + P.dv;
            if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2)
                xbase[p] = (P.sx * RAW_T + (v0 - 1 - (Y0 - 2))) * 8 + lx;
            if (MODE == 0 && !a.first_init && inimg) {
                const long pix = (long)(Y * M + P.sy) * a.wM + (X * M + P.sx);
                s_avg[p] = a.out[pix];
                s_e2[p] = a.out[pix + whM];
                s_tac[p] = a.out[pix + 2 * whM];
                s_int[p] = a.out[pix + 3 * whM];
                s_prv[p] = a.prev[pix];
            }
        }
    }

    const int wI = 2 * w, hI = 2 * h;
    const long whI = (long)wI * hI, wh = (long)w * h;
    const int kxo = 2 * X0 - a.G0, kyo = 2 * Y0 - a.G0;
    const int planesz = a.PP * a.PH;
    const int base = ly * a.PP + lx;
    const int pw = a.TW >> 1;
    const int nU = 4 * a.PH * pw;

    // fill plan of the U tile (plane-major, as in k_bic_rgc) and the class's tap tables
    for (int e = tid; e < nU; e += nthr) {
        const int pl = e / (a.PH * pw);
        const int r = (e - pl * a.PH * pw) / pw;
        const int c = e - pl * a.PH * pw - r * pw;
        const int gy = 2 * r + (pl >> 1), gx = 2 * c + (pl & 1);
        lPlan[e] = (pl * planesz + r * a.PP + c) | (gy << 15) | (gx << 23);
    }
    const int tcls0 = PHS[0].first_tap;
    if (!TSM) {
        const int tcls1 = PHS[MM - 1].first_tap + PHS[MM - 1].ntaps;
        for (int e = tid; e < tcls1 - tcls0 + TAP_SPARE; e += nthr) lT[e] = taps[tcls0 + e];
    }
    // x weights and column offset of each column phase sx (phase (0, sx) of the class: both
    // depend on sx only)
    for (int e = tid; e < M; e += nthr) {
        const Phase P = PHS[e];
        lXW[e] = make_float4(P.wx[0], P.wx[1], P.wx[2], P.wx[3]);
        lXD[e] = P.du + 1;
    }
    __syncthreads();

    // The bicubic intensity (liveSRRF.cl:361) sums each of the 4 rows along x first, then
    // the 4 row sums along y.  A row sum depends on the tile row, the macro-pixel column lx
    // and the column phase sx only, so the block forms each (sx, tile row, lx) sum of frames
    // f and f+1 once per frame pair, in the same order, and every phase of that column
    // phase reads four of them.
    auto xpass = [&]() {
        const int nxs = M * XS_PER_SX;
        for (int e = tid; e < nxs; e += nthr) {
            const int sx = e / XS_PER_SX;
            const int rr = e - sx * XS_PER_SX;
            const float4 wv = lXW[sx];
            const float2 *src = lR + (rr >> 3) * RAW_P + (rr & 7) + lXD[sx];
            const float2 v0 = src[0], v1 = src[1], v2 = src[2], v3 = src[3];
            float pa = 0.0f, pb = 0.0f;
            pa = fmaf(v0.x, wv.x, pa);
            pb = fmaf(v0.y, wv.x, pb);
            pa = fmaf(v1.x, wv.y, pa);
            pb = fmaf(v1.y, wv.y, pb);
            pa = fmaf(v2.x, wv.z, pa);
            pb = fmaf(v2.y, wv.z, pb);
            pa = fmaf(v3.x, wv.w, pa);
            pb = fmaf(v3.y, wv.w, pb);
            lXS[e] = make_float2(pa, pb);
        }
    };

    // frames f and f+1 (f+1 only when `two`) into the LDS tiles
    auto fill = [&](int f, bool two) {
        const float2 *Ua = a.U + whI * f;
        for (int e = tid; e < nU; e += nthr) {
            const int pk = lPlan[e];
            const int gy = (pk >> 15) & 255, gx = (pk >> 23) & 255;
            const int kx = kxo + gx, ky = kyo + gy;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (gy < a.TH && gx < a.TW && kx >= 0 && kx < wI && ky >= 0 && ky < hI) {
                const long g = (long)ky * wI + kx;
                const float2 ua = Ua[g];
                const float2 ub = two ? Ua[g + whI] : make_float2(0.0f, 0.0f);
                v = make_float4(ua.x, ua.y, ub.x, ub.y);
            }
            lU[pk & 0x7fff] = v;
        }
        const float *Pa = a.px + wh * f;
        for (int e = tid; e < RAW_T * RAW_T; e += nthr) {
            const int r = e / RAW_T, c = e - r * RAW_T;
            const long g = (long)min(max(Y0 - 2 + r, 0), h - 1) * w + min(max(X0 - 2 + c, 0), w - 1);
            lR[r * RAW_P + c] = make_float2(Pa[g], two ? Pa[g + wh] : 0.0f);
        }
    };

    // register staging (STG > 0; the launch guarantees nU <= STG * nthr and
    // RAW_T^2 <= nthr).  A thread stages the same tile elements for every frame pair, so
    // their offsets within a frame (-1: outside the image, the element stays 0) and their
    // LDS slots are formed once.
    float4 su[STG > 0 ? STG : 1];
    float2 sr = make_float2(0.0f, 0.0f);
    int sgo[STG > 0 ? STG : 1], sdst[STG > 0 ? STG : 1];
    int rgo = -1;
#pragma unroll
    for (int k = 0; k < STG; k++) {
        const int e = tid + k * nthr;
        sgo[k] = -1;
        sdst[k] = -1;
        if (e < nU) {
            const int pk = lPlan[e];
            const int gy = (pk >> 15) & 255, gx = (pk >> 23) & 255;
            const int kx = kxo + gx, ky = kyo + gy;
            sdst[k] = pk & 0x7fff;
            if (gy < a.TH && gx < a.TW && kx >= 0 && kx < wI && ky >= 0 && ky < hI)
                sgo[k] = ky * wI + kx;
        }
    }
    if (STG && tid < RAW_T * RAW_T) {
        const int r = tid / RAW_T, c = tid - r * RAW_T;
        rgo = min(max(Y0 - 2 + r, 0), h - 1) * w + min(max(X0 - 2 + c, 0), w - 1);
    }
    auto prefetch = [&](int f, bool two) {
        const float2 *Ua = a.U + whI * f;
#pragma unroll
        for (int k = 0; k < STG; k++) {
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (sgo[k] >= 0) {
                const float2 ua = Ua[sgo[k]];
                const float2 ub = two ? Ua[sgo[k] + whI] : make_float2(0.0f, 0.0f);
                v = make_float4(ua.x, ua.y, ub.x, ub.y);
            }
            su[k] = v;
        }
        if (rgo >= 0) {
            const float *Pa = a.px + wh * f;
            sr = make_float2(Pa[rgo], two ? Pa[rgo + wh] : 0.0f);
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (int k = 0; k < STG; k++)
            if (sdst[k] >= 0) lU[sdst[k]] = su[k];
        if (rgo >= 0) {
            const int r = tid / RAW_T, c = tid - r * RAW_T;
            lR[r * RAW_P + c] = sr;
        }
    };

    fill(a.frame0, a.nframes > 1);
    __syncthreads();
    xpass();
    __syncthreads();

    for (int fi = 0; fi < a.nframes; fi += 2) {
        const bool two = fi + 1 < a.nframes;
        const bool first = (fi == 0) && a.first_init;
        if (STG && fi + 2 < a.nframes) prefetch(a.frame0 + fi + 2, fi + 3 < a.nframes);
        // (a.zero is 0 at run time; it keeps the per-phase table reads inside the frame loop)
        const int zf = a.zero * fi;
        const Phase *PHf = PHS + zf;

#pragma unroll PH_UNROLL
        for (int p = 0; p < PPW; p++) {
            const int ph = wave + p * a.nw;
            if (ph < MM) {
                const Phase P = PHf[ph];
                float acc0 = 0.0f, acc1 = 0.0f;
                const Tap *Tl = TSM ? taps + P.first_tap : lT + (P.first_tap - tcls0);
                // (one chunk of 4 taps per iteration: two in flight exceed the register
                // budget of 4 waves per SIMD; with TSM the next chunk's taps are loaded
                // into SGPRs during this one -- every phase's list is a multiple of 4 long
                // and the table ends in TAP_SPARE zero taps, so t + 7 stays inside it)
                Tap n0, n1, n2, n3;
                if (TSM) {
                    n0 = Tl[0];
                    n1 = Tl[1];
                    n2 = Tl[2];
                    n3 = Tl[3];
                }
                // (from the LDS tables the loop stops at the even tap count: the zero-weight
                // padding beyond it adds -0 to each sum)
                const int n4 = TSM ? P.ntaps : (P.ntaps2 & ~3);
                int t = 0;
#pragma unroll 1
                for (; t < n4; t += 4) {
                    Tap c0, c1, c2, c3;
                    if (TSM) {
                        c0 = n0;
                        c1 = n1;
                        c2 = n2;
                        c3 = n3;
                        n0 = Tl[t + 4];
                        n1 = Tl[t + 5];
                        n2 = Tl[t + 6];
                        n3 = Tl[t + 7];
                    } else {
                        c0 = Tl[t];
                        c1 = Tl[t + 1];
                        c2 = Tl[t + 2];
                        c3 = Tl[t + 3];
                    }
                    const float4 u0 = lU[base + c0.off];
                    const float4 u1 = lU[base + c1.off];
                    const float4 u2 = lU[base + c2.off];
                    const float4 u3 = lU[base + c3.off];
                    acc0 = tap_acc(make_float2(u0.x, u0.y), c0, acc0);
                    acc1 = tap_acc(make_float2(u0.z, u0.w), c0, acc1);
                    acc0 = tap_acc(make_float2(u1.x, u1.y), c1, acc0);
                    acc1 = tap_acc(make_float2(u1.z, u1.w), c1, acc1);
                    acc0 = tap_acc(make_float2(u2.x, u2.y), c2, acc0);
                    acc1 = tap_acc(make_float2(u2.z, u2.w), c2, acc1);
                    acc0 = tap_acc(make_float2(u3.x, u3.y), c3, acc0);
                    acc1 = tap_acc(make_float2(u3.z, u3.w), c3, acc1);
                }
                if (!TSM && (P.ntaps2 & 2)) {
                    const Tap c0 = Tl[t], c1 = Tl[t + 1];
                    const float4 u0 = lU[base + c0.off];
                    const float4 u1 = lU[base + c1.off];
                    acc0 = tap_acc(make_float2(u0.x, u0.y), c0, acc0);
                    acc1 = tap_acc(make_float2(u0.z, u0.w), c0, acc1);
                    acc0 = tap_acc(make_float2(u1.x, u1.y), c1, acc0);
                    acc1 = tap_acc(make_float2(u1.z, u1.w), c1, acc1);
                }
                const float C0a = (P.wsum + acc0) * invw[p];
                const float C0b = (P.wsum + acc1) * invw[p];

                // interpolated intensity at (xM/M - 0.5, yM/M - 0.5) (liveSRRF.cl:361)
                float va, vb;
                if (xbase[p] >= 0) {
                    const float2 *xs = lXS + xbase[p];
                    float qa = 0.0f, qb = 0.0f;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const float2 r2 = xs[j * 8];
                        qa = fmaf(r2.x, P.wy[j], qa);
                        qb = fmaf(r2.y, P.wy[j], qb);
                    }
                    va = qa;
                    vb = qb;
                } else {
                    // border pixel: the reference's bilinear-extrapolation regime
                    const int xM = (X + zf) * M + P.sx, yM = (Y + zf) * M + P.sy;
                    const float xf = (float)xM / (float)M - 0.5f;
                    const float yf = (float)yM / (float)M - 0.5f;
                    const float *lRf = reinterpret_cast<const float *>(lR);
                    va = interp_tile_s<2>(lRf, RAW_P, X0 - 2, Y0 - 2, w, h, xf, yf);
                    vb = interp_tile_s<2>(lRf + 1, RAW_P, X0 - 2, Y0 - 2, w, h, xf, yf);
                }

                if (MODE == 1) {
                    if (inimg) {
                        const long e = (((long)(a.frame0 + fi) * MM + ph) * h + Y) * w + X;
                        a.emitC[e] = C0a;
                        a.emitV[e] = va;
                        if (two) {
                            a.emitC[e + (long)MM * wh] = C0b;
                            a.emitV[e + (long)MM * wh] = vb;
                        }
                    }
                    continue;
                }
                cumulate<INTW>(s_avg[p], s_e2[p], s_tac[p], s_int[p], s_prv[p], C0a, va, first,
                               a);
                if (two)
                    cumulate<INTW>(s_avg[p], s_e2[p], s_tac[p], s_int[p], s_prv[p], C0b, vb,
                                   false, a);
            }
        }
        if (fi + 2 < a.nframes) {
            __syncthreads();   // every wave is done with this pair's tiles
            if (STG)
                commit();
            else
                fill(a.frame0 + fi + 2, fi + 3 < a.nframes);
            __syncthreads();
            xpass();
            __syncthreads();
        }
    }

    if (MODE == 0 && inimg) {
#pragma unroll
        for (int p = 0; p < PPW; p++) {
            const int ph = wave + p * a.nw;
            if (ph >= MM) continue;
            const Phase P = PHS[ph];
            const long pix = (long)(Y * M + P.sy) * a.wM + (X * M + P.sx);
            a.out[pix] = s_avg[p];
            a.out[pix + whM] = s_e2[p];
            a.out[pix + 2 * whM] = s_tac[p];
            a.out[pix + 3 * whM] = s_int[p];
            a.prev[pix] = s_prv[p];
        }
    }
}

// ---------------------------------------------------------------------------------------
// Four frames per pass (mode 0, tap tables in LDS, parity-plane pitch 24).  The U tile of
// frames f, f+1 and the tile of frames f+2, f+3 sit QUAD_UPL float4 elements apart, a
// compile-time distance, so one tap read and one address serve four samples: the second
// tile's read takes the first's address with an immediate offset.  The raw tiles hold the
// four frames as float4 (pitch 24: the two rows of a 16-lane pass fall in opposite bank
// halves).  One set of barriers per four frames.  Taps are walked two at a time (every
// phase's gated count is rounded up to an even number); each frame's sum takes the taps in
// order and the cumulants are updated frame by frame, so results equal k_bic_rgc2's bit
// for bit.  MODE 1 emits each frame's normalised CGLH and intensity instead (the emit order
// of k_bic_rgc2's MODE 1).
// ---------------------------------------------------------------------------------------
constexpr int QUAD_PP = 24;                    // parity-plane pitch the kernel assumes
constexpr int QUAD_UPL = 4 * QUAD_PP * 14;     // float4 slots of one tile (parity planes of
                                               // up to 14 rows, TW <= 28)
constexpr int RAW_P4 = 24;                     // raw-tile pitch (float4 elements)

template <int PPW, bool INTW, int STG, int MODE>
__global__ void __launch_bounds__(1024)
__attribute__((amdgpu_waves_per_eu(4))) k_bic_rgc4(const BicRgcArgs a,
                                                   const Tap *__restrict__ taps,
                                                   const Phase *__restrict__ phases,
                                                   const int *__restrict__ tapk)
{
    extern __shared__ __align__(16) unsigned char smem[];
    float4 *lU = reinterpret_cast<float4 *>(smem);                 // [2][QUAD_UPL]
    float4 *lR = lU + 2 * QUAD_UPL;                                  // [RAW_T][RAW_P4]
    float4 *lXW = lR + RAW_T * RAW_P4;                               // [M]
    float4 *lXS = lXW + a.M;                                         // [M][RAW_T][8]
    int *lXD = reinterpret_cast<int *>(lXS + a.M * XS_PER_SX);      // [M]
    Tap *lT = static_cast<Tap *>(__builtin_assume_aligned(lXD + XS_MAX_M, 16));
    int *lPlan = reinterpret_cast<int *>(lT + a.max_class_taps + TAP_SPARE);
    constexpr int PH_UNROLL = PPW <= 4 ? PPW : 1;

    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lx = lane & 7, ly = lane >> 3;
    const int lin = xcd_tile_1d(a.tiles_x * a.tiles_y);
    const int tX = lin % a.tiles_x, tY = lin / a.tiles_x;
    const int X0 = tX << 3, Y0 = tY << 3;
    const int X = X0 + lx, Y = Y0 + ly;
    const int M = a.M, MM = M * M;
    const int w = a.w, h = a.h;
    const Phase *PHS = phases + (axis_class(Y0, a.xs_log2) * a.ncls_x +
                                   axis_class(X0, a.xs_log2)) * MM;
    const bool interior = (2 * X0 + a.R2lo >= 1) && (2 * (X0 + 7) + 1 + a.R2hi <= 2 * w - 1) &&
                          (2 * Y0 + a.R2lo >= 1) && (2 * (Y0 + 7) + 1 + a.R2hi <= 2 * h - 1);
    const bool inimg = (X < w) && (Y < h);
    const long whM = (long)a.wM * a.hM;

    float s_avg[PPW], s_e2[PPW], s_tac[PPW], s_int[PPW], s_prv[PPW];
    float invw[PPW];
    int xbase[PPW];
#pragma unroll
    for (int p = 0; p < PPW; p++) {
        const int ph = wave + p * a.nw;
        invw[p] = 0.0f;
        xbase[p] = -1;
        s_avg[p] = s_e2[p] = s_tac[p] = s_int[p] = s_prv[p] = 0.0f;
        if (ph < MM) {
            const Phase P = PHS[ph];
            float ws = P.wsum;
            if (!interior) {
                ws = 0.0f;
                for (int t = 0; t < P.ntaps; t++) {
                    const int k = tapk[P.first_tap + t];
                    const int kx = 2 * X + (k & 255) - 128;
                    const int ky = 2 * Y + ((k >> 8) & 255) - 128;
                    if (kx >= 1 && kx <= 2 * w - 1 && ky >= 1 && ky <= 2 * h - 1)
                        ws += taps[P.first_tap + t].W;
                }
            }
            invw[p] = 1.0f / ws;
            const int u0 = X + P.du, v0 = Y + P.dv;
            if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2)
                xbase[p] = (P.sx * RAW_T + (v0 - 1 - (Y0 - 2))) * 8 + lx;
            if (MODE == 0 && !a.first_init && inimg) {
                const long pix = (long)(Y * M + P.sy) * a.wM + (X * M + P.sx);
                s_avg[p] = a.out[pix];
                s_e2[p] = a.out[pix + whM];
                s_tac[p] = a.out[pix + 2 * whM];
                s_int[p] = a.out[pix + 3 * whM];
                s_prv[p] = a.prev[pix];
            }
        }
    }

    const int wI = 2 * w, hI = 2 * h;
    const long whI = (long)wI * hI, wh = (long)w * h;
    const int kxo = 2 * X0 - a.G0, kyo = 2 * Y0 - a.G0;
    const int planesz = a.PP * a.PH;
    const int base = ly * a.PP + lx;
    const int pw = a.TW >> 1;
    const int nU = 4 * a.PH * pw;

    for (int e = tid; e < nU; e += nthr) {
        const int pl = e / (a.PH * pw);
        const int r = (e - pl * a.PH * pw) / pw;
        const int c = e - pl * a.PH * pw - r * pw;
        const int gy = 2 * r + (pl >> 1), gx = 2 * c + (pl & 1);
        lPlan[e] = (pl * planesz + r * a.PP + c) | (gy << 15) | (gx << 23);
    }
    const int tcls0 = PHS[0].first_tap;
    {
        const int tcls1 = PHS[MM - 1].first_tap + PHS[MM - 1].ntaps;
        for (int e = tid; e < tcls1 - tcls0 + TAP_SPARE; e += nthr) lT[e] = taps[tcls0 + e];
    }
    for (int e = tid; e < M; e += nthr) {
        const Phase P = PHS[e];
        lXW[e] = make_float4(P.wx[0], P.wx[1], P.wx[2], P.wx[3]);
        lXD[e] = P.du + 1;
    }
    __syncthreads();

    // x row sums of the bicubic intensity (liveSRRF.cl:361) for the four frames, as in
    // k_bic_rgc2
    auto xpass = [&]() {
        const int nxs = M * XS_PER_SX;
        for (int e = tid; e < nxs; e += nthr) {
            const int sx = e / XS_PER_SX;
            const int rr = e - sx * XS_PER_SX;
            const float4 wv = lXW[sx];
            const float4 *src = lR + (rr >> 3) * RAW_P4 + (rr & 7) + lXD[sx];
            const float4 v0 = src[0], v1 = src[1], v2 = src[2], v3 = src[3];
            float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            q.x = fmaf(v0.x, wv.x, q.x);
            q.y = fmaf(v0.y, wv.x, q.y);
            q.z = fmaf(v0.z, wv.x, q.z);
            q.w = fmaf(v0.w, wv.x, q.w);
            q.x = fmaf(v1.x, wv.y, q.x);
            q.y = fmaf(v1.y, wv.y, q.y);
            q.z = fmaf(v1.z, wv.y, q.z);
            q.w = fmaf(v1.w, wv.y, q.w);
            q.x = fmaf(v2.x, wv.z, q.x);
            q.y = fmaf(v2.y, wv.z, q.y);
            q.z = fmaf(v2.z, wv.z, q.z);
            q.w = fmaf(v2.w, wv.z, q.w);
            q.x = fmaf(v3.x, wv.w, q.x);
            q.y = fmaf(v3.y, wv.w, q.y);
            q.z = fmaf(v3.z, wv.w, q.z);
            q.w = fmaf(v3.w, wv.w, q.w);
            lXS[e] = q;
        }
    };

    // frames f .. f+nf-1 (nf <= 4; missing frames read as 0) into the LDS tiles
    auto fill = [&](int f, int nf) {
        const float2 *Ua = a.U + whI * f;
        for (int e = tid; e < nU; e += nthr) {
            const int pk = lPlan[e];
            const int gy = (pk >> 15) & 255, gx = (pk >> 23) & 255;
            const int kx = kxo + gx, ky = kyo + gy;
            float2 u[4] = {};
            if (gy < a.TH && gx < a.TW && kx >= 0 && kx < wI && ky >= 0 && ky < hI) {
                const long g = (long)ky * wI + kx;
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (j < nf) u[j] = Ua[g + j * whI];
            }
            lU[pk & 0x7fff] = make_float4(u[0].x, u[0].y, u[1].x, u[1].y);
            lU[QUAD_UPL + (pk & 0x7fff)] = make_float4(u[2].x, u[2].y, u[3].x, u[3].y);
        }
        const float *Pa = a.px + wh * f;
        for (int e = tid; e < RAW_T * RAW_T; e += nthr) {
            const int r = e / RAW_T, c = e - r * RAW_T;
            const long g = (long)min(max(Y0 - 2 + r, 0), h - 1) * w + min(max(X0 - 2 + c, 0), w - 1);
            float v[4] = {};
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (j < nf) v[j] = Pa[g + j * wh];
            lR[r * RAW_P4 + c] = make_float4(v[0], v[1], v[2], v[3]);
        }
    };

    // register staging of the next four frames (STG > 0; the launch guarantees
    // nU <= STG * nthr and RAW_T^2 <= nthr)
    float4 su[STG > 0 ? STG : 1], su2[STG > 0 ? STG : 1];
    float4 sr = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    int sgo[STG > 0 ? STG : 1], sdst[STG > 0 ? STG : 1];
    int rgo = -1;
#pragma unroll
    for (int k = 0; k < STG; k++) {
        const int e = tid + k * nthr;
        sgo[k] = -1;
        sdst[k] = -1;
        if (e < nU) {
            const int pk = lPlan[e];
            const int gy = (pk >> 15) & 255, gx = (pk >> 23) & 255;
            const int kx = kxo + gx, ky = kyo + gy;
            sdst[k] = pk & 0x7fff;
            if (gy < a.TH && gx < a.TW && kx >= 0 && kx < wI && ky >= 0 && ky < hI)
                sgo[k] = ky * wI + kx;
        }
    }
    if (STG && tid < RAW_T * RAW_T) {
        const int r = tid / RAW_T, c = tid - r * RAW_T;
        rgo = min(max(Y0 - 2 + r, 0), h - 1) * w + min(max(X0 - 2 + c, 0), w - 1);
    }
    auto prefetch = [&](int f, int nf) {
        const float2 *Ua = a.U + whI * f;
#pragma unroll
        for (int k = 0; k < STG; k++) {
            float2 u[4] = {};
            if (sgo[k] >= 0) {
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (j < nf) u[j] = Ua[sgo[k] + j * whI];
            }
            su[k] = make_float4(u[0].x, u[0].y, u[1].x, u[1].y);
            su2[k] = make_float4(u[2].x, u[2].y, u[3].x, u[3].y);
        }
        if (rgo >= 0) {
            const float *Pa = a.px + wh * f;
            float v[4] = {};
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (j < nf) v[j] = Pa[rgo + j * wh];
            sr = make_float4(v[0], v[1], v[2], v[3]);
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (int k = 0; k < STG; k++)
            if (sdst[k] >= 0) {
                lU[sdst[k]] = su[k];
                lU[QUAD_UPL + sdst[k]] = su2[k];
            }
        if (rgo >= 0) {
            const int r = tid / RAW_T, c = tid - r * RAW_T;
            lR[r * RAW_P4 + c] = sr;
        }
    };

    fill(a.frame0, min(4, a.nframes));
    __syncthreads();
    xpass();
    __syncthreads();

    for (int fi = 0; fi < a.nframes; fi += 4) {
        const int nf = min(4, a.nframes - fi);
        const bool first = (fi == 0) && a.first_init;
        if (STG && fi + 4 < a.nframes) prefetch(a.frame0 + fi + 4, min(4, a.nframes - fi - 4));
        const int zf = a.zero * fi;
        const Phase *PHf = PHS + zf;

#pragma unroll PH_UNROLL
        for (int p = 0; p < PPW; p++) {
            const int ph = wave + p * a.nw;
            if (ph < MM) {
                const Phase P = PHf[ph];
                float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, acc3 = 0.0f;
                const Tap *Tl = lT + (P.first_tap - tcls0);
                // (the loop stops at the even gated count: the zero-weight padding beyond it
                // adds -0 to each sum)
#pragma unroll 1
                for (int t = 0; t < P.ntaps2; t += 2) {
                    const Tap c0 = Tl[t], c1 = Tl[t + 1];
                    const float4 *p0 = lU + base + c0.off;
                    const float4 *p1 = lU + base + c1.off;
                    const float4 a0 = p0[0], b0 = p0[QUAD_UPL];
                    const float4 a1 = p1[0], b1 = p1[QUAD_UPL];
                    acc0 = tap_acc(make_float2(a0.x, a0.y), c0, acc0);
                    acc1 = tap_acc(make_float2(a0.z, a0.w), c0, acc1);
                    acc2 = tap_acc(make_float2(b0.x, b0.y), c0, acc2);
                    acc3 = tap_acc(make_float2(b0.z, b0.w), c0, acc3);
                    acc0 = tap_acc(make_float2(a1.x, a1.y), c1, acc0);
                    acc1 = tap_acc(make_float2(a1.z, a1.w), c1, acc1);
                    acc2 = tap_acc(make_float2(b1.x, b1.y), c1, acc2);
                    acc3 = tap_acc(make_float2(b1.z, b1.w), c1, acc3);
                }
                const float C0[4] = {(P.wsum + acc0) * invw[p], (P.wsum + acc1) * invw[p],
                                     (P.wsum + acc2) * invw[p], (P.wsum + acc3) * invw[p]};
                float v[4];
                if (xbase[p] >= 0) {
                    const float4 *xs = lXS + xbase[p];
                    float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const float4 r4 = xs[j * 8];
                        q.x = fmaf(r4.x, P.wy[j], q.x);
                        q.y = fmaf(r4.y, P.wy[j], q.y);
                        q.z = fmaf(r4.z, P.wy[j], q.z);
                        q.w = fmaf(r4.w, P.wy[j], q.w);
                    }
                    v[0] = q.x;
                    v[1] = q.y;
                    v[2] = q.z;
                    v[3] = q.w;
                } else {
                    const int xM = (X + zf) * M + P.sx, yM = (Y + zf) * M + P.sy;
                    const float xf = (float)xM / (float)M - 0.5f;
                    const float yf = (float)yM / (float)M - 0.5f;
                    const float *lRf = reinterpret_cast<const float *>(lR);
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        v[j] = interp_tile_s<4>(lRf + j, RAW_P4, X0 - 2, Y0 - 2, w, h, xf, yf);
                }
                if (MODE == 1) {
                    if (inimg) {
                        const long e = (((long)(a.frame0 + fi) * MM + ph) * h + Y) * w + X;
#pragma unroll
                        for (int j = 0; j < 4; j++)
                            if (j < nf) {
                                a.emitC[e + j * (long)MM * wh] = C0[j];
                                a.emitV[e + j * (long)MM * wh] = v[j];
                            }
                    }
                    continue;
                }
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (j < nf)
                        cumulate<INTW>(s_avg[p], s_e2[p], s_tac[p], s_int[p], s_prv[p], C0[j],
                                       v[j], first && j == 0, a);
            }
        }
        if (fi + 4 < a.nframes) {
            __syncthreads();   // every wave is done with these frames' tiles
            if (STG)
                commit();
            else
                fill(a.frame0 + fi + 4, min(4, a.nframes - fi - 4));
            __syncthreads();
            xpass();
            __syncthreads();
        }
    }

    if (MODE == 0 && inimg) {
#pragma unroll
        for (int p = 0; p < PPW; p++) {
            const int ph = wave + p * a.nw;
            if (ph >= MM) continue;
            const Phase P = PHS[ph];
            const long pix = (long)(Y * M + P.sy) * a.wM + (X * M + P.sx);
            a.out[pix] = s_avg[p];
            a.out[pix + whM] = s_e2[p];
            a.out[pix + 2 * whM] = s_tac[p];
            a.out[pix + 3 * whM] = s_int[p];
            a.prev[pix] = s_prv[p];
        }
    }
}

// ---------------------------------------------------------------------------------------
// k_bic_emit_multi: the parameter sweep's emit pass for up to SWEEP_NFG fwhm values of one
// magnification at once.  A sample's angular term h = (1 - |sin|)[cos < 0] - 1 depends on
// the tap's offset and direction only; the fwhm sets the weights and the gate.  The union
// tables (build_bic_geometry_multi) list the widest gate's taps in the reference's order,
// each with the weight of every fwhm of the group (0 outside its gate), so one h per tap
// and frame feeds NFG accumulators: fwhm f's sum sees its own taps in its own order and
// adds W * h = +0 for the others, which leaves it bit for bit what k_bic_rgc2 (mode 1)
// computes from fwhm f's own tables.  The intensity is interpolated once per phase and
// frame.  Emits CGLH of fwhm f to emitC + f * cstride and the intensity to emitV, both
// phase-major [frame][M*M][h][w].  One block: an 8x8 macro-pixel tile and PG = nw * PPW
// consecutive sub-pixel phases (grid.y: phase groups), whose taps alone are staged in LDS;
// two frames per pass.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ float tap_h_ab(float2 u, float ta, float tb)
{
    float c2 = fmaf(u.x, -ta, -1.0f);
    c2 = fmaf(u.y, -tb, c2);
    float s = u.y * ta;
    s = fmaf(u.x, -tb, s);
    return __builtin_amdgcn_fmed3f(-1.0f, -fabsf(s), c2);
}

size_t bic_emit_multi_lds_bytes(const BicRgcArgs &a, int pg_taps)
{
    return sizeof(float4) * 4 * (size_t)a.PP * a.PH + sizeof(float2) * RAW_SZ +
           sizeof(TapM) * ((size_t)pg_taps + TAP_SPARE) +
           sizeof(int) * (size_t)(4 * a.PH * (a.TW >> 1));
}

template <int PPW, int NFG>
__global__ void __launch_bounds__(1024) k_bic_emit_multi(const BicRgcArgs a,
                                                         const TapM *__restrict__ taps,
                                                         const PhaseM *__restrict__ phases,
                                                         const int *__restrict__ tapk,
                                                         int pg_taps, long cstride)
{
    extern __shared__ __align__(16) unsigned char smem[];
    float4 *lU = reinterpret_cast<float4 *>(smem);
    const int usz = 4 * a.PP * a.PH;
    float2 *lR = reinterpret_cast<float2 *>(lU + usz);
    float4 *lT = reinterpret_cast<float4 *>(lR + RAW_SZ);   // TapM as two float4 each
    int *lPlan = reinterpret_cast<int *>(lT + 2 * (pg_taps + TAP_SPARE));

    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lx = lane & 7, ly = lane >> 3;
    const int lin = xcd_tile_1d(a.tiles_x * a.tiles_y);
    const int tX = lin % a.tiles_x, tY = lin / a.tiles_x;
    const int X0 = tX << 3, Y0 = tY << 3;
    const int X = X0 + lx, Y = Y0 + ly;
    const int M = a.M, MM = M * M;
    const int w = a.w, h = a.h;
    const int ph0 = blockIdx.y * (a.nw * PPW);
    const int phl = min(ph0 + a.nw * PPW, MM) - 1;
    const PhaseM *PHS = phases + (axis_class(Y0, a.xs_log2) * a.ncls_x +
                                  axis_class(X0, a.xs_log2)) * MM;
    const int tb0 = PHS[ph0].first_tap;
    const int tb1 = PHS[phl].first_tap + PHS[phl].ntaps;
    const bool interior = (2 * X0 + a.R2lo >= 1) && (2 * (X0 + 7) + 1 + a.R2hi <= 2 * w - 1) &&
                          (2 * Y0 + a.R2lo >= 1) && (2 * (Y0 + 7) + 1 + a.R2hi <= 2 * h - 1);
    const bool inimg = (X < w) && (Y < h);
    const long wh = (long)w * h;

    float invw[PPW][NFG];
    int ibase[PPW];
#pragma unroll
    for (int p = 0; p < PPW; p++) {
        const int ph = ph0 + wave * PPW + p;
        ibase[p] = -1;
#pragma unroll
        for (int f = 0; f < NFG; f++) invw[p][f] = 0.0f;
        if (ph < MM) {
            const PhaseM &P = PHS[ph];
            float ws[NFG];
#pragma unroll
            for (int f = 0; f < NFG; f++) ws[f] = P.wsum[f];
            if (!interior) {
                // border: sum only the taps inside the 0 < v < size gates, reference order
#pragma unroll
                for (int f = 0; f < NFG; f++) ws[f] = 0.0f;
                for (int t = 0; t < P.ntaps; t++) {
                    const int k = tapk[P.first_tap + t];
                    const int kx = 2 * X + (k & 255) - 128;
                    const int ky = 2 * Y + ((k >> 8) & 255) - 128;
                    if (kx >= 1 && kx <= 2 * w - 1 && ky >= 1 && ky <= 2 * h - 1) {
#pragma unroll
                        for (int f = 0; f < NFG; f++) ws[f] += taps[P.first_tap + t].W[f];
                    }
                }
            }
#pragma unroll
            for (int f = 0; f < NFG; f++) invw[p][f] = 1.0f / ws[f];
            const int u0 = X + P.du, v0 = Y + P.dv;
            if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2)
                ibase[p] = (v0 - 1 - (Y0 - 2)) * RAW_P + (u0 - 1 - (X0 - 2));
        }
    }

    const int wI = 2 * w, hI = 2 * h;
    const long whI = (long)wI * hI;
    const int kxo = 2 * X0 - a.G0, kyo = 2 * Y0 - a.G0;
    const int planesz = a.PP * a.PH;
    const int base = ly * a.PP + lx;
    const int pw = a.TW >> 1;
    const int nU = 4 * a.PH * pw;
    for (int e = tid; e < nU; e += nthr) {
        const int pl = e / (a.PH * pw);
        const int r = (e - pl * a.PH * pw) / pw;
        const int c = e - pl * a.PH * pw - r * pw;
        const int gy = 2 * r + (pl >> 1), gx = 2 * c + (pl & 1);
        lPlan[e] = (pl * planesz + r * a.PP + c) | (gy << 15) | (gx << 23);
    }
    {
        const float4 *T4 = reinterpret_cast<const float4 *>(taps + tb0);
        for (int e = tid; e < 2 * (tb1 - tb0 + TAP_SPARE); e += nthr) lT[e] = T4[e];
    }
    __syncthreads();

    auto fill = [&](int f, bool two) {
        const float2 *Ua = a.U + whI * f;
        for (int e = tid; e < nU; e += nthr) {
            const int pk = lPlan[e];
            const int gy = (pk >> 15) & 255, gx = (pk >> 23) & 255;
            const int kx = kxo + gx, ky = kyo + gy;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (gy < a.TH && gx < a.TW && kx >= 0 && kx < wI && ky >= 0 && ky < hI) {
                const long g = (long)ky * wI + kx;
                const float2 ua = Ua[g];
                const float2 ub = two ? Ua[g + whI] : make_float2(0.0f, 0.0f);
                v = make_float4(ua.x, ua.y, ub.x, ub.y);
            }
            lU[pk & 0x7fff] = v;
        }
        const float *Pa = a.px + wh * f;
        for (int e = tid; e < RAW_T * RAW_T; e += nthr) {
            const int r = e / RAW_T, c = e - r * RAW_T;
            const long g = (long)min(max(Y0 - 2 + r, 0), h - 1) * w + min(max(X0 - 2 + c, 0), w - 1);
            lR[r * RAW_P + c] = make_float2(Pa[g], two ? Pa[g + wh] : 0.0f);
        }
    };

    fill(a.frame0, a.nframes > 1);
    __syncthreads();

    for (int fi = 0; fi < a.nframes; fi += 2) {
        const bool two = fi + 1 < a.nframes;
#pragma unroll
        for (int p = 0; p < PPW; p++) {
            const int ph = ph0 + wave * PPW + p;
            if (ph >= MM) continue;
            const PhaseM &P = PHS[ph];
            float acc[NFG][2];
#pragma unroll
            for (int f = 0; f < NFG; f++) acc[f][0] = acc[f][1] = 0.0f;
            const float4 *Tl = lT + 2 * (P.first_tap - tb0);
#pragma unroll 1
            for (int t = 0; t < P.ntaps; t += 4) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const float4 th = Tl[2 * (t + k)];
                    const float4 tw = Tl[2 * (t + k) + 1];
                    const float4 u = lU[base + __float_as_int(th.x)];
                    const float h0 = tap_h_ab(make_float2(u.x, u.y), th.y, th.z);
                    const float h1 = tap_h_ab(make_float2(u.z, u.w), th.y, th.z);
                    const float W[5] = {th.w, tw.x, tw.y, tw.z, tw.w};
#pragma unroll
                    for (int f = 0; f < NFG; f++) {
                        acc[f][0] = fmaf(W[f], h0, acc[f][0]);
                        acc[f][1] = fmaf(W[f], h1, acc[f][1]);
                    }
                }
            }
            // interpolated intensity at (xM/M - 0.5, yM/M - 0.5) (liveSRRF.cl:361)
            float va, vb;
            if (ibase[p] >= 0) {
                const float2 *row = lR + ibase[p];
                float qa = 0.0f, qb = 0.0f;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    float pa = 0.0f, pb = 0.0f;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const float2 r2 = row[j * RAW_P + i];
                        pa = fmaf(r2.x, P.wx[i], pa);
                        pb = fmaf(r2.y, P.wx[i], pb);
                    }
                    qa = fmaf(pa, P.wy[j], qa);
                    qb = fmaf(pb, P.wy[j], qb);
                }
                va = qa;
                vb = qb;
            } else {
                const int xM = X * M + P.sx, yM = Y * M + P.sy;
                const float xf = (float)xM / (float)M - 0.5f;
                const float yf = (float)yM / (float)M - 0.5f;
                const float *lRf = reinterpret_cast<const float *>(lR);
                va = interp_tile_s<2>(lRf, RAW_P, X0 - 2, Y0 - 2, w, h, xf, yf);
                vb = interp_tile_s<2>(lRf + 1, RAW_P, X0 - 2, Y0 - 2, w, h, xf, yf);
            }
            if (inimg) {
                const long e = (((long)(a.frame0 + fi) * MM + ph) * h + Y) * w + X;
                a.emitV[e] = va;
                if (two) a.emitV[e + (long)MM * wh] = vb;
#pragma unroll
                for (int f = 0; f < NFG; f++) {
                    float *Cf = a.emitC + cstride * f;
                    Cf[e] = (P.wsum[f] + acc[f][0]) * invw[p][f];
                    if (two) Cf[e + (long)MM * wh] = (P.wsum[f] + acc[f][1]) * invw[p][f];
                }
            }
        }
        if (fi + 2 < a.nframes) {
            __syncthreads();   // every wave is done with this pair's tiles
            fill(a.frame0 + fi + 2, fi + 3 < a.nframes);
            __syncthreads();
        }
    }
}

template <int PPW>
static hipError_t launch_emit_multi_p(const BicRgcArgs &a, const TapM *taps, const PhaseM *ph,
                                      const int *tapk, int nfg, int pg_taps, long cstride,
                                      hipStream_t s)
{
    const int MM = a.M * a.M;
    const dim3 grid(a.tiles_x * a.tiles_y, (MM + a.nw * PPW - 1) / (a.nw * PPW));
    const size_t lds = bic_emit_multi_lds_bytes(a, pg_taps);
    switch (nfg) {
    case 1: hipLaunchKernelGGL((k_bic_emit_multi<PPW, 1>), grid, dim3(64 * a.nw), lds, s, a, taps, ph, tapk, pg_taps, cstride); break;
    case 2: hipLaunchKernelGGL((k_bic_emit_multi<PPW, 2>), grid, dim3(64 * a.nw), lds, s, a, taps, ph, tapk, pg_taps, cstride); break;
    case 3: hipLaunchKernelGGL((k_bic_emit_multi<PPW, 3>), grid, dim3(64 * a.nw), lds, s, a, taps, ph, tapk, pg_taps, cstride); break;
    case 4: hipLaunchKernelGGL((k_bic_emit_multi<PPW, 4>), grid, dim3(64 * a.nw), lds, s, a, taps, ph, tapk, pg_taps, cstride); break;
    default: hipLaunchKernelGGL((k_bic_emit_multi<PPW, 5>), grid, dim3(64 * a.nw), lds, s, a, taps, ph, tapk, pg_taps, cstride); break;
    }
    return hipGetLastError();
}

hipError_t launch_bic_emit_multi(const BicRgcArgs &a, int ppw, const TapM *taps,
                                 const PhaseM *phases, const int *tapk, int nfg, int pg_taps,
                                 long cstride, hipStream_t s)
{
    if (a.nframes <= 0) return hipSuccess;
    if (nfg < 1 || nfg > SWEEP_NFG || a.nw < 1 || a.nw > 16) return hipErrorInvalidValue;
    return ppw == 2 ? launch_emit_multi_p<2>(a, taps, phases, tapk, nfg, pg_taps, cstride, s)
                    : launch_emit_multi_p<1>(a, taps, phases, tapk, nfg, pg_taps, cstride, s);
}

size_t bic_pair_lds_bytes(const BicRgcArgs &a, bool tsm)
{
    return sizeof(float4) * 4 * (size_t)a.PP * a.PH + sizeof(float2) * RAW_SZ +
           (sizeof(float4) + sizeof(float2) * XS_PER_SX) * (size_t)a.M +
           sizeof(int) * XS_MAX_M +
           (tsm ? 0 : sizeof(Tap) * ((size_t)a.max_class_taps + TAP_SPARE)) +
           sizeof(int) * (size_t)(4 * a.PH * (a.TW >> 1));
}

size_t bic_quad_lds_bytes(const BicRgcArgs &a)
{
    return sizeof(float4) * (2 * (size_t)QUAD_UPL + RAW_T * RAW_P4) +
           (sizeof(float4) + sizeof(float4) * XS_PER_SX) * (size_t)a.M +
           sizeof(int) * XS_MAX_M + sizeof(Tap) * ((size_t)a.max_class_taps + TAP_SPARE) +
           sizeof(int) * (size_t)(4 * a.PH * (a.TW >> 1));
}

bool bic_quad_fits(const BicRgcArgs &a)
{
    return a.PP == QUAD_PP && 4 * a.PP * a.PH <= QUAD_UPL && a.M <= XS_MAX_M &&
           bic_quad_lds_bytes(a) <= 80 * 1024;
}

template <int PPW, bool INTW, int MODE>
static hipError_t launch_quad(const BicRgcArgs &a, hipStream_t s)
{
    // register staging needs every tile element within 2 per thread
    const int nthr = 64 * a.nw, nU = 4 * a.PH * (a.TW >> 1);
    const size_t lds = bic_quad_lds_bytes(a);
    if (!a.pair_direct && nthr >= RAW_T * RAW_T && nU <= 2 * nthr)
        hipLaunchKernelGGL((k_bic_rgc4<PPW, INTW, 2, MODE>), dim3(a.tiles_x * a.tiles_y),
                           dim3(nthr), lds, s, a, a.taps, a.phases, a.tapk);
    else
        hipLaunchKernelGGL((k_bic_rgc4<PPW, INTW, 0, MODE>), dim3(a.tiles_x * a.tiles_y),
                           dim3(nthr), lds, s, a, a.taps, a.phases, a.tapk);
    return hipGetLastError();
}

template <int PPW>
static hipError_t launch_quad_m(const BicRgcArgs &a, hipStream_t s)
{
    if (a.mode == 1) return launch_quad<PPW, false, 1>(a, s);
    return a.int_weighting ? launch_quad<PPW, true, 0>(a, s) : launch_quad<PPW, false, 0>(a, s);
}

// taps_in_lds 0: the tap tables are read with scalar loads (no LDS copy)
template <int PPW, int MODE, bool INTW, int STG>
static hipError_t launch_pair_s(const BicRgcArgs &a, hipStream_t s)
{
    if (!a.taps_in_lds)
        hipLaunchKernelGGL((k_bic_rgc2<PPW, MODE, INTW, true, STG>),
                           dim3(a.tiles_x * a.tiles_y), dim3(64 * a.nw),
                           bic_pair_lds_bytes(a, true), s, a, a.taps, a.phases, a.tapk);
    else
        hipLaunchKernelGGL((k_bic_rgc2<PPW, MODE, INTW, false, STG>),
                           dim3(a.tiles_x * a.tiles_y), dim3(64 * a.nw),
                           bic_pair_lds_bytes(a, false), s, a, a.taps, a.phases, a.tapk);
    return hipGetLastError();
}

template <int PPW, int MODE, bool INTW>
static hipError_t launch_pair_m(const BicRgcArgs &a, hipStream_t s)
{
    // register staging (the default) needs every tile element within 2 per thread
    const int nthr = 64 * a.nw, nU = 4 * a.PH * (a.TW >> 1);
    if (!a.pair_direct && nthr >= RAW_T * RAW_T && nU <= 2 * nthr)
        return launch_pair_s<PPW, MODE, INTW, 2>(a, s);
    return launch_pair_s<PPW, MODE, INTW, 0>(a, s);
}

template <int PPW>
static hipError_t launch_pair(const BicRgcArgs &a, hipStream_t s)
{
    if (a.mode == 1) return launch_pair_m<PPW, 1, false>(a, s);
    return a.int_weighting ? launch_pair_m<PPW, 0, true>(a, s)
                           : launch_pair_m<PPW, 0, false>(a, s);
}

template <int PPW, bool TLDS, int MODE>
static hipError_t launch_ppw_m(const BicRgcArgs &a, hipStream_t s)
{
    const size_t one = sizeof(float2) * 4 * a.PP * a.PH + sizeof(float) * RAW_SZ;
    const size_t tl = (TLDS ? (sizeof(Tap) + (MODE == 2 ? sizeof(float) : 0)) *
                                  ((size_t)a.max_class_taps + TAP_SPARE)
                            : 0) +
                      sizeof(int) * (size_t)(4 * a.PH * (a.TW >> 1));   // fill plan
    const int nthr = 64 * a.nw;
    const int nU = 4 * a.PH * (a.TW >> 1);
    const int need = (nU + nthr - 1) / nthr;
    const dim3 grid(a.tiles_x * a.tiles_y), block(nthr);
    // register staging of the next frame's tiles costs VGPRs; where it would push the
    // kernel below its occupancy target the tiles are filled directly after the compute
    if (a.stage && nthr >= RAW_T * RAW_T && need <= 2)
        hipLaunchKernelGGL((k_bic_rgc<PPW, 2, TLDS, MODE>), grid, block, 2 * one + tl, s, a,
                           a.taps, a.phases, a.tapk);
    else if (a.stage && nthr >= RAW_T * RAW_T && need <= 4)
        hipLaunchKernelGGL((k_bic_rgc<PPW, 4, TLDS, MODE>), grid, block, 2 * one + tl, s, a,
                           a.taps, a.phases, a.tapk);
    else
        hipLaunchKernelGGL((k_bic_rgc<PPW, 0, TLDS, MODE>), grid, block, one + tl, s, a,
                           a.taps, a.phases, a.tapk);
    return hipGetLastError();
}

template <int PPW, bool TLDS>
static hipError_t launch_ppw_t(const BicRgcArgs &a, hipStream_t s)
{
    switch (a.mode) {
    case 1: return launch_ppw_m<PPW, TLDS, 1>(a, s);
    case 2: return launch_ppw_m<PPW, TLDS, 2>(a, s);
    default: return launch_ppw_m<PPW, TLDS, 0>(a, s);
    }
}

template <int PPW>
static hipError_t launch_ppw(const BicRgcArgs &a, hipStream_t s)
{
    return a.taps_in_lds ? launch_ppw_t<PPW, true>(a, s) : launch_ppw_t<PPW, false>(a, s);
}

hipError_t launch_bic_rgc(const BicRgcArgs &a, int ppw, hipStream_t s, const char **kname)
{
    if (a.nframes <= 0) return hipSuccess;
    const bool quad = a.frames_per_pass == 4 && a.mode != 2 && ppw <= 4 && a.tac_prev_scaled &&
                      a.taps_in_lds && bic_quad_fits(a);
    const bool pair = !quad && a.frames_per_pass >= 2 && a.mode != 2 && ppw <= 4 &&
                      a.tac_prev_scaled && bic_pair_lds_bytes(a, !a.taps_in_lds) <= 64 * 1024;
    if (kname) *kname = quad ? "k_bic_rgc4" : pair ? "k_bic_rgc2" : "k_bic_rgc";
    if (quad) {
        switch (ppw) {
        case 1: return launch_quad_m<1>(a, s);
        case 2: return launch_quad_m<2>(a, s);
        case 3: return launch_quad_m<3>(a, s);
        default: return launch_quad_m<4>(a, s);
        }
    }
    if (pair) {
        switch (ppw) {
        case 1: return launch_pair<1>(a, s);
        case 2: return launch_pair<2>(a, s);
        case 3: return launch_pair<3>(a, s);
        default: return launch_pair<4>(a, s);
        }
    }
    switch (ppw) {
    case 1: return launch_ppw<1>(a, s);
    case 2: return launch_ppw<2>(a, s);
    case 3: return launch_ppw<3>(a, s);
    case 4: return launch_ppw<4>(a, s);
    case 5: return launch_ppw<5>(a, s);
    case 6: return launch_ppw<6>(a, s);
    case 7: return launch_ppw<7>(a, s);
    case 8: return launch_ppw<8>(a, s);
    default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------------------
// Generic per-pixel RGC (drift shifts, any M): the reference loop structure of
// liveSRRF.cl:257-393 evaluated on the unit tap gradients.
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_generic_rgc(const GenericRgcArgs a)
{
    const long offset = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long whM = (long)a.wM * a.hM;
    if (offset >= whM) return;
    const int yM = (int)(offset / a.wM);
    const int xM = (int)(offset - (long)yM * a.wM);
    const int w = a.w, h = a.h, M = a.M;
    const int wI = 2 * w, hI = 2 * h;
    const long whI = (long)wI * hI, wh = (long)w * h;
    const int jlo = -(int)(2.0f * a.radius);
    const int jhi = (int)(2.0f * a.radius + 1.0f);
    float s0, s1, s2, s3, s4;
    if (a.count0 == 0) {
        s0 = s1 = s2 = s3 = s4 = 0.0f;
    } else {
        s0 = a.out[offset];
        s1 = a.out[offset + whM];
        s2 = a.out[offset + 2 * whM];
        s3 = a.out[offset + 3 * whM];
        s4 = a.prev[offset];
    }
    for (int fi = 0; fi < a.nframes; fi++) {
        const int c0 = a.count0 + fi;
        const int f = a.frame0 + fi;
        const float shiftX = a.drift[c0];
        const float shiftY = a.drift[c0 + a.nFrameForSRRF];
        const float xc = (float)((xM + 0.5) / M + (double)shiftX);
        const float yc = (float)((yM + 0.5) / M + (double)shiftY);
        const float2 *Uf = a.U + whI * f;
        float CGLH = 0.0f, sumW = 0.0f;
        const int bxi = (int)(2.0f * xc), byi = (int)(2.0f * yc);
        for (int j = jlo; j <= jhi; j++) {
            const float vy = ((float)byi + (float)j) / 2.0f;
            if (!(vy > 0.0f && vy < (float)h)) continue;
            for (int i = jlo; i <= jhi; i++) {
                const float vx = ((float)bxi + (float)i) / 2.0f;
                if (!(vx > 0.0f && vx < (float)w)) continue;
                const float dx = vx - xc, dy = vy - yc;
                const float d = sqrtf(dx * dx + dy * dy);
                if (!(d != 0.0f && d <= a.tSO)) continue;
                float dw = d * expf(-(d * d) / a.tSS);
                dw = dw * dw * dw * dw;
                sumW += dw;
                const int n = bxi + i, m = byi + j;   // 2*vx, 2*vy
                const float2 u = Uf[(long)m * wI + n];
                const float gdotr = u.x * dx + u.y * dy;
                if (gdotr < 0.0f) {
                    float Dk = fabsf(u.y * dx - u.x * dy);
                    Dk = 1.0f - Dk / d;
                    CGLH += Dk * dw;
                }
            }
        }
        CGLH /= sumW;
        CGLH = apply_sens(CGLH, a.sens, a.sens_int);
        const float v = interp_global(a.px + wh * f, w, h, ((float)xM) / (float)M + shiftX - 0.5f,
                                      ((float)yM) / (float)M + shiftY - 0.5f);
        const float t = a.int_weighting ? v * CGLH : CGLH;
        const float e2 = a.int_weighting ? (t * v) * CGLH : CGLH * CGLH;
        if (c0 == 0) {
            s0 = t * a.rn;
            s1 = e2 * a.rn;
            s2 = 0.0f;
            s3 = v * a.rn;
        } else {
            s0 = s0 + t * a.rn;
            s1 = s1 + e2 * a.rn;
            s2 = s2 + (t * s4) * a.rn1;
            s3 = s3 + v * a.rn;
        }
        s4 = a.tac_prev_scaled ? t * a.rn1 : t;
    }
    a.out[offset] = s0;
    a.out[offset + whM] = s1;
    a.out[offset + 2 * whM] = s2;
    a.out[offset + 3 * whM] = s3;
    a.prev[offset] = s4;
}

hipError_t launch_generic_rgc(const GenericRgcArgs &a, hipStream_t s)
{
    if (a.nframes <= 0) return hipSuccess;
    const long n = (long)a.wM * a.hM;
    hipLaunchKernelGGL(k_generic_rgc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace esrrf
