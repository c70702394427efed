// k_3d.hip -- 3-D (multi-focus microscopy) eSRRF: resources/live3DSRRF.cl and the montage
// split of LiveSRRF_CL.reshapeImageStack3D (LiveSRRF_CL.java:963-994).
//
// Per load:
//   k3d_reshape   montage cell -> plane, (p - background) / intensity scaling  (:983-986)
//   k3d_align     rotate about the plane centre by -theta, shift (live3DSRRF.cl:204-224)
//   k3d_grad      2-point gradients Gx, Gy, Gz, edge-clamped (live3DSRRF.cl:229-279)
//   k3d_unit      the three staggered reads of the lateral 2x interpolated gradients the
//                 RGC makes at a sample point (live3DSRRF.cl:282-320, :389-391)
//                     T(n, m, z) = (GxInt[z][m][n+1], GyInt[z][m+1][n], GzInt[z][m][n])
//                 folded into one unit 3-vector per point of the 2x grid.
// Per launch of frames:
//   k3d_lateral   lateral Catmull-Rom intensity of every plane at every magnified (x, y)
//                 (the four per-voxel plane samples of live3DSRRF.cl:441-447 come from it)
//   k3d_rgc       3-D radial gradient convergence + axial cubic + temporal cumulants
//                 (live3DSRRF.cl:325-498), cumulants in registers across frames.  Frames
//                 without drift share one geometry, evaluated once per chunk of FC frames.
#include <algorithm>
#include <cstdlib>

#include "device_interp.h"
#include "esrrf_internal.h"
#include "mode3d_internal.h"

namespace esrrf {

static inline unsigned grid1(long n, int bs = 256)
{
    long g = (n + bs - 1) / bs;
    if (g > 65535L * 64) g = 65535L * 64;
    return (unsigned)(g < 1 ? 1 : g);
}

// ---------------------------------------------------------------------------------------
__global__ void k3d_reshape(const void *__restrict__ src, int dtype, float *__restrict__ planes,
                            int W, int H, int w, int h, int nP, const int *__restrict__ cell,
                            const float *__restrict__ ic, float bg, long n)
{
    const long wh = (long)w * h, whd = wh * nP;
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < n;
         o += (long)gridDim.x * blockDim.x) {
        const long f = o / whd;
        const int z = (int)((o - f * whd) / wh);
        const long r = o - f * whd - (long)z * wh;
        const int yy = (int)(r / w), xx = (int)(r - (long)yy * w);
        const long s = (long)W * H * f + (long)(cell[nP + z] + yy) * W + (cell[z] + xx);
        const float p = dtype == ESRRF_U16 ? (float)((const unsigned short *)src)[s]
                                           : ((const float *)src)[s];
        planes[o] = (p - bg) / ic[z];
    }
}

__global__ void k3d_align(const float *__restrict__ planes, float *__restrict__ aligned, int w,
                          int h, int nP, const float *__restrict__ st, long n)
{
    const long wh = (long)w * h, whd = wh * nP;
    const float cw = (float)(w - 1) / 2.0f, ch = (float)(h - 1) / 2.0f;
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < n;
         o += (long)gridDim.x * blockDim.x) {
        const long f = o / whd;
        const int z = (int)((o - f * whd) / wh);
        const long r = o - f * whd - (long)z * wh;
        const int y = (int)(r / w), x = (int)(r - (long)y * w);
        const float ct = cosf(st[z]), sn = sinf(st[z]);
        const float xr = (float)x - cw, yr = (float)y - ch;
        float xa = __fmul_rn(xr, ct);
        xa = __fsub_rn(xa, __fmul_rn(yr, sn));
        xa = __fadd_rn(xa, cw);
        xa = __fsub_rn(xa, st[z + nP]);
        float ya = __fmul_rn(xr, sn);
        ya = __fadd_rn(ya, __fmul_rn(yr, ct));
        ya = __fadd_rn(ya, ch);
        ya = __fsub_rn(ya, st[z + 2 * nP]);
        aligned[o] = interp_global(planes + whd * f + wh * z, w, h, xa, ya);
    }
}

// g: [3][nF*nP][h][w]
__global__ void k3d_grad(const float *__restrict__ px, float *__restrict__ g, int w, int h,
                         int nP, long n)
{
    const long wh = (long)w * h, whd = wh * nP;
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < n;
         o += (long)gridDim.x * blockDim.x) {
        const long f = o / whd;
        const int z1 = (int)((o - f * whd) / wh);
        const long r = o - f * whd - (long)z1 * wh;
        const int y1 = (int)(r / w), x1 = (int)(r - (long)y1 * w);
        const int x0 = max(x1 - 1, 0), y0 = max(y1 - 1, 0), z0 = max(z1 - 1, 0);
        const float p = px[o];
        const long base = whd * f;
        g[o] = p - px[base + wh * z1 + (long)y1 * w + x0];
        g[n + o] = p - px[base + wh * z1 + (long)y0 * w + x1];
        g[2 * n + o] = p - px[base + wh * z0 + (long)y1 * w + x1];
    }
}

__device__ __forceinline__ float4 unit3_or_zero(float tx, float ty, float tz)
{
    const float m = fmaxf(fmaxf(fabsf(tx), fabsf(ty)), fabsf(tz));
    if (!(m > 0.0f) || !(m <= 3.4e38f)) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float sx = tx / m, sy = ty / m, sz = tz / m;
    const float r = rsqrtf(sx * sx + sy * sy + sz * sz);
    return make_float4(sx * r, sy * r, sz * r, 0.0f);
}

// GxInt-style value of plane `a` at the 2x-grid point (X/2, Y/2) (getInterpolatedValue,
// live3DSRRF.cl:30-140).  At integer and half-integer arguments the Catmull-Rom weights are
// exactly (0, 1, 0, 0) and (-1, 9, 9, -1)/16 (cubic(0) = 1, cubic(1) = cubic(2) = 0,
// cubic(1/2) = 9/16, cubic(3/2) = -1/16), so they are taken by parity instead of evaluated;
// the products are summed in the same order, rows along x first (interp_global).
__device__ __forceinline__ float half_grid_value(const float *__restrict__ a, int w, int h,
                                                 int X, int Y)
{
    const int u0 = X >> 1, v0 = Y >> 1;
    if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2) {
        const float c0 = -0.0625f, c1 = 0.5625f;
        const float wx0 = (X & 1) ? c0 : 0.0f, wx1 = (X & 1) ? c1 : 1.0f;
        const float wx2 = (X & 1) ? c1 : 0.0f, wx3 = (X & 1) ? c0 : 0.0f;
        const float wy[4] = {(Y & 1) ? c0 : 0.0f, (Y & 1) ? c1 : 1.0f, (Y & 1) ? c1 : 0.0f,
                             (Y & 1) ? c0 : 0.0f};
        const float *r = a + (v0 - 1) * w + (u0 - 1);
        float q = 0.0f;
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            float p = 0.0f;
            p = fmaf(r[0], wx0, p);
            p = fmaf(r[1], wx1, p);
            p = fmaf(r[2], wx2, p);
            p = fmaf(r[3], wx3, p);
            q = fmaf(p, wy[jj], q);
            r += w;
        }
        return q;
    }
    return interp_global(a, w, h, (float)X / 2.0f, (float)Y / 2.0f);
}

// U: [nF*nP][2h][2w]; the last column and row of the 2x grid are never sampled (gates).
// Block: 32 x 32 points of the 2x grid of one plane (planes over grid.z).  The three
// gradient tiles the block reaches (image columns and rows kx0/2 - 1 .. kx0/2 + 18) are
// staged in LDS.  A 4-tap sum along x of a gradient row depends only on the row and the
// point's x, so the block forms each (row, x) sum once and every point combines four of them
// along y, in half_grid_value's order; border points take its extrapolation branch.
constexpr int U3_T = 32;   // points per block side
constexpr int U3_G = 20;   // gradient tile side

__global__ void __launch_bounds__(256) k3d_unit(const float *__restrict__ g,
                                                float4 *__restrict__ U, int w, int h, long npl)
{
    __shared__ float lg[3][U3_G * U3_G];
    __shared__ float rs[3][U3_G * U3_T];   // row sums [plane c][tile row][point x]
    const int wI = 2 * w, hI = 2 * h;
    const long wh = (long)w * h, gn = npl * wh, whI = (long)wI * hI;
    const int kx0 = blockIdx.x * U3_T, ky0 = blockIdx.y * U3_T;
    const int gx0 = (kx0 >> 1) - 1, gy0 = (ky0 >> 1) - 1;
    const int tid = threadIdx.x;
    const float c0 = -0.0625f, c1 = 0.5625f;
    for (long pl = blockIdx.z; pl < npl; pl += gridDim.z) {
        const float *G[3] = {g + wh * pl, g + gn + wh * pl, g + 2 * gn + wh * pl};
        __syncthreads();   // the previous plane's tiles are no longer read
        for (int e = tid; e < 3 * U3_G * U3_G; e += 256) {
            const int c = e / (U3_G * U3_G), rr = e - c * U3_G * U3_G;
            const int r = rr / U3_G, q = rr - r * U3_G;
            const int x = gx0 + q, y = gy0 + r;
            lg[c][rr] = (x >= 0 && x < w && y >= 0 && y < h)
                            ? g[c * gn + wh * pl + (long)y * w + x] : 0.0f;
        }
        __syncthreads();
        // row sums along x: Gx at X = kx + 1, Gy and Gz at X = kx
        for (int e = tid; e < 3 * U3_G * U3_T; e += 256) {
            const int c = e / (U3_G * U3_T), rr = e - c * U3_G * U3_T;
            const int r = rr / U3_T, q = rr - r * U3_T;
            const int X = kx0 + q + (c == 0 ? 1 : 0), u0 = X >> 1;
            const float wx0 = (X & 1) ? c0 : 0.0f, wx1 = (X & 1) ? c1 : 1.0f;
            const float wx2 = (X & 1) ? c1 : 0.0f, wx3 = (X & 1) ? c0 : 0.0f;
            const int col = u0 - 1 - gx0;
            float pp = 0.0f;
            if (col >= 0 && col + 3 < U3_G) {
                const float *t = &lg[c][r * U3_G + col];
                pp = fmaf(t[0], wx0, pp);
                pp = fmaf(t[1], wx1, pp);
                pp = fmaf(t[2], wx2, pp);
                pp = fmaf(t[3], wx3, pp);
            }
            rs[c][rr] = pp;
        }
        __syncthreads();
        float4 *Up = U + pl * whI;
#pragma unroll 1
        for (int o = 0; o < (U3_T * U3_T) / 256; o++) {
            const int e = tid + o * 256;
            const int q = e & (U3_T - 1), yy = e / U3_T;
            const int nn = kx0 + q, m = ky0 + yy;
            if (nn >= wI || m >= hI) continue;
            float4 u = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (nn <= wI - 2 && m <= hI - 2) {
                float t3[3];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const int X = nn + (c == 0 ? 1 : 0), Y = m + (c == 1 ? 1 : 0);
                    const int u0 = X >> 1, v0 = Y >> 1;
                    if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2) {
                        const float wy0 = (Y & 1) ? c0 : 0.0f, wy1 = (Y & 1) ? c1 : 1.0f;
                        const float wy2 = (Y & 1) ? c1 : 0.0f, wy3 = (Y & 1) ? c0 : 0.0f;
                        const float *cs = &rs[c][(v0 - 1 - gy0) * U3_T + q];
                        float qq = 0.0f;
                        qq = fmaf(cs[0], wy0, qq);
                        qq = fmaf(cs[U3_T], wy1, qq);
                        qq = fmaf(cs[2 * U3_T], wy2, qq);
                        qq = fmaf(cs[3 * U3_T], wy3, qq);
                        t3[c] = qq;
                    } else {
                        t3[c] = interp_global(G[c], w, h, (float)X / 2.0f, (float)Y / 2.0f);
                    }
                }
                u = unit3_or_zero(t3[0], t3[1], t3[2]);
            }
            Up[(long)m * wI + nn] = u;
        }
    }
}

// GxInt, GyInt, GzInt of frame 0, every plane: out [nP][3][2h][2w] (LiveSRRF_CL.java:855-872)
__global__ void k3d_gint(const float *__restrict__ g, float *__restrict__ out, int w, int h,
                         int nP, long gn)
{
    const int wI = 2 * w, hI = 2 * h;
    const long whI = (long)wI * hI, wh = (long)w * h;
    const long n = (long)nP * 3 * whI;
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < n;
         o += (long)gridDim.x * blockDim.x) {
        const int z = (int)(o / (3 * whI));
        const int c = (int)((o - (long)z * 3 * whI) / whI);
        const long r = o - (long)z * 3 * whI - (long)c * whI;
        const int y = (int)(r / wI), x = (int)(r - (long)y * wI);
        out[o] = interp_global(g + c * gn + wh * z, w, h, (float)x / 2.0f, (float)y / 2.0f);
    }
}

// L: [nframes][nP][hM][wM].  Block: 64 x 16 magnified pixels of one (frame, plane)
// (frame-planes over grid.z).  The raw pixels the block reaches are staged in LDS: sample
// positions x = xM/M + dX - 0.5 increase with xM, so the columns floor(x(x0)) - 1 ..
// floor(x(x0 + 63)) + 2 (at most 68 for M >= 1) and the rows likewise (at most 20) cover
// every bicubic stencil of the block.  A thread keeps its column's Catmull-Rom weights for
// its four rows; weights, products and sums are interp_global's (same arguments, same order).
constexpr int LAT_X = 64, LAT_Y = 16, LAT_TW = LAT_X + 4, LAT_TH = LAT_Y + 4;

__global__ void __launch_bounds__(256)
k3d_lateral(const float *__restrict__ px, float *__restrict__ L, int w, int h, int nP, int M,
            int frame0, int nframes, const float *__restrict__ drift, int count0, int nFor)
{
    __shared__ float tile[LAT_TH * LAT_TW];
    const int wM = w * M, hM = h * M;
    const long whM = (long)wM * hM, wh = (long)w * h;
    const int x0 = blockIdx.x * LAT_X, y0 = blockIdx.y * LAT_Y;
    const int tid = threadIdx.x, tx = tid & (LAT_X - 1), rg = tid >> 6;
    const int xM = x0 + tx;
    const int nfz = nframes * nP;
#pragma unroll 1
    for (int fz = blockIdx.z; fz < nfz; fz += gridDim.z) {
        const int fi = fz / nP, z = fz - fi * nP;
        const float dX = drift[count0 + fi], dY = drift[count0 + fi + nFor];
        const float *P = px + wh * ((long)(frame0 + fi) * nP + z);
        const float xs0 = __fsub_rn(__fadd_rn((float)x0 / (float)M, dX), 0.5f);
        const float ys0 = __fsub_rn(__fadd_rn((float)y0 / (float)M, dY), 0.5f);
        const int ux0 = (int)floorf(xs0) - 1, vy0 = (int)floorf(ys0) - 1;
        __syncthreads();   // the previous frame-plane's tile is no longer read
        for (int e = tid; e < LAT_TH * LAT_TW; e += 256) {
            const int r = e / LAT_TW, c = e - r * LAT_TW;
            const int x = ux0 + c, y = vy0 + r;
            tile[e] = (x >= 0 && x < w && y >= 0 && y < h) ? P[(long)y * w + x] : 0.0f;
        }
        __syncthreads();
        if (xM >= wM) continue;
        const float xi = __fsub_rn(__fadd_rn((float)xM / (float)M, dX), 0.5f);
        const int u0 = (int)floorf(xi);
        const bool inx = u0 > 0 && u0 < w - 2;
        float wx[4];
#pragma unroll
        for (int i = 0; i < 4; i++) wx[i] = cubic_w(xi - (float)(u0 - 1 + i));
        float *Lo = L + (long)fz * whM + xM;
#pragma unroll 1
        for (int k = 0; k < LAT_Y / 4; k++) {
            const int yM = y0 + rg + 4 * k;
            if (yM >= hM) break;
            const float yi = __fsub_rn(__fadd_rn((float)yM / (float)M, dY), 0.5f);
            const int v0 = (int)floorf(yi);
            float q = 0.0f;
            if (inx && v0 > 0 && v0 < h - 2) {
                const float *t = &tile[(v0 - 1 - vy0) * LAT_TW + (u0 - 1 - ux0)];
#pragma unroll
                for (int j = 0; j <= 3; j++) {
                    float p = 0.0f;
                    p = fmaf(t[0], wx[0], p);
                    p = fmaf(t[1], wx[1], p);
                    p = fmaf(t[2], wx[2], p);
                    p = fmaf(t[3], wx[3], p);
                    q = fmaf(p, cubic_w(yi - (float)(v0 - 1 + j)), q);
                    t += LAT_TW;
                }
            } else {
                q = interp_global(P, w, h, xi, yi);
            }
            Lo[(long)yM * wM] = q;
        }
    }
}

__device__ __forceinline__ float apply_sens3(float C, float sens, int sens_int)
{
    if (!(C >= 0.0f)) return 0.0f;
    switch (sens_int) {
    case 1: return C;
    case 2: return C * C;
    case 3: return C * C * C;
    case 4: { const float c2 = C * C; return c2 * c2; }
    default: return powf(C, sens);
    }
}

__device__ __forceinline__ float cross_mag3(float a1, float a2, float a3, float b1, float b2,
                                            float b3)
{
    const float G1 = a2 * b3 - a3 * b2;
    const float G2 = a3 * b1 - a1 * b3;
    const float G3 = a1 * b2 - a2 * b1;
    return sqrtf(G1 * G1 + G2 * G2 + G3 * G3);
}

// One thread per output voxel; block 64 x 4 voxels of one magnified z-slice.
template <int FC>
__global__ void __launch_bounds__(256) k3d_rgc(const Rgc3dArgs a)
{
    const int xM = blockIdx.x * 64 + threadIdx.x;
    const int yM = blockIdx.y * 4 + threadIdx.y;
    const int zM = blockIdx.z;
    if (xM >= a.wM || yM >= a.hM) return;
    const int w = a.w, h = a.h, nP = a.nP, M = a.M;
    const long whM = (long)a.wM * a.hM;
    const long whdM = whM * nP * a.Mz;
    const long vox = whM * zM + (long)yM * a.wM + xM;
    const long whI = 4L * w * h;
    const int wI = 2 * w;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f, s4 = 0.0f;
    if (a.count0 > 0) {
        s0 = a.out[vox];
        s1 = a.out[vox + whdM];
        s2 = a.out[vox + 2 * whdM];
        s3 = a.out[vox + 3 * whdM];
        s4 = a.prev[vox];
    }
    const float zc = (float)((zM + 0.5) / a.Mz);
    int z0;
    const float zcof = zc - 0.5f;
    if (zcof < 0.0f) z0 = 0;
    else if (zcof >= (float)(nP - 1)) z0 = nP - 2;
    else z0 = (int)floorf(zcof);
    const int zm1 = max(z0 - 1, 0), zp2 = min(z0 + 1, nP - 1);
    const float z01 = zc - ((float)z0 + 0.5f);
    const int jlo = -(int)(2.0f * a.radius), jhi = (int)(2.0f * a.radius + 1.0f);
    const int klo = -(int)a.radius, khi = (int)a.radius + 1;
    const float nFf = (float)a.nFor, nF1 = (float)(a.nFor - 1);

    for (int f0 = 0; f0 < a.nframes; f0 += FC) {
        const int nf = min(FC, a.nframes - f0);
        const int c00 = a.count0 + f0;
        const float driftX = a.drift[c00], driftY = a.drift[c00 + a.nFor];
        const float xc = (float)((xM + 0.5) / M + (double)driftX);
        const float yc = (float)((yM + 0.5) / M + (double)driftY);
        float acc[FC];
#pragma unroll
        for (int q = 0; q < FC; q++) acc[q] = 0.0f;
        float sumW = 0.0f;
        const float4 *Ub = a.U + (long)(a.frame0 + f0) * nP * whI;
        const float bx = (float)((int)(2.0f * xc)), by = (float)((int)(2.0f * yc));
        for (int j = jlo; j <= jhi; j++) {
            const float vy = (by + (float)j) / 2.0f;
            if (!(vy > 0.0f && vy < (float)h)) continue;
            const int m = (int)(2.0 * ((double)vy - 0.5));
            for (int i = jlo; i <= jhi; i++) {
                const float vx = (bx + (float)i) / 2.0f;
                if (!(vx > 0.0f && vx < (float)w)) continue;
                const int nn = (int)(2.0 * ((double)vx - 0.5));
                const float dx = vx - xc, dy = vy - yc;
                for (int k = klo; k <= khi; k++) {
                    const float vz = (float)((int)zc) + (float)k;
                    if (!(vz > 0.0f && vz < (float)nP)) continue;
                    const float dz = vz - zc;
                    float d2 = __fmul_rn(dx, dx);
                    d2 = __fadd_rn(d2, __fmul_rn(dy, dy));
                    d2 = __fadd_rn(d2, __fmul_rn(dz, dz));
                    const float d = sqrtf(d2);
                    if (!(d != 0.0f && d <= a.tSO)) continue;
                    float dw = d * expf(-(d * d) / a.tSS);
                    dw = dw * dw * dw * dw;
                    sumW += dw;
                    const float invd = 1.0f / d;
                    const long idx = whI * (int)vz + (long)m * wI + nn;
#pragma unroll
                    for (int q = 0; q < FC; q++) {
                        if (q < nf) {
                            const float4 u = Ub[(long)q * nP * whI + idx];
                            const float gd = u.x * dx + u.y * dy + u.z * dz;
                            const float Dk = 1.0f - cross_mag3(dx, dy, dz, u.x, u.y, u.z) * invd;
                            acc[q] += gd < 0.0f ? Dk * dw : 0.0f;
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < FC; q++) {
            if (q < nf) {
                const int fi = f0 + q;
                float C = acc[q] / sumW;
                C = apply_sens3(C, a.sens, a.sens_int);
                const float *Lf = a.L + (long)fi * nP * whM + (long)yM * a.wM + xM;
                const float v_1 = Lf[whM * zm1], v0 = Lf[whM * z0];
                const float v1 = Lf[whM * (z0 + 1)], v2 = Lf[whM * zp2];
                // axial cubic (live3DSRRF.cl:448-452), left-to-right as written
                float aa = __fmul_rn(-0.5f, v_1);
                aa = __fadd_rn(aa, __fmul_rn(1.5f, v0));
                aa = __fsub_rn(aa, __fmul_rn(1.5f, v1));
                aa = __fadd_rn(aa, __fmul_rn(0.5f, v2));
                float bb = __fsub_rn(v_1, __fmul_rn(2.5f, v0));
                bb = __fadd_rn(bb, __fmul_rn(2.0f, v1));
                bb = __fsub_rn(bb, __fmul_rn(0.5f, v2));
                const float cc = __fadd_rn(__fmul_rn(-0.5f, v_1), __fmul_rn(0.5f, v1));
                float v = __fmul_rn(__fmul_rn(__fmul_rn(aa, z01), z01), z01);
                v = __fadd_rn(v, __fmul_rn(__fmul_rn(bb, z01), z01));
                v = __fadd_rn(v, __fmul_rn(cc, z01));
                v = __fadd_rn(v, v0);
                const float t = a.int_weighting ? v * C : C;
                const float e2 = a.int_weighting ? ((t * v) * C) : C * C;
                if (c00 + q == 0) {
                    s0 = t / nFf;
                    s1 = e2 / nFf;
                    s2 = 0.0f;
                    s3 = v / nFf;
                } else {
                    s0 = s0 + t / nFf;
                    s1 = s1 + e2 / nFf;
                    s2 = s2 + (t * s4) / nF1;
                    s3 = s3 + v / nFf;
                }
                s4 = t;   // single 1/(n-1) in TAC2 (live3DSRRF.cl:472)
            }
        }
    }
    a.out[vox] = s0;
    a.out[vox + whdM] = s1;
    a.out[vox + 2 * whdM] = s2;
    a.out[vox + 3 * whdM] = s3;
    a.prev[vox] = s4;
}

// ---------------------------------------------------------------------------------------
// Frames without drift: the geometry of a voxel (sample points, distances, weights, unit
// directions e = d/|d|) is the same for every frame, so it is computed once per chunk of FC
// frames and each frame pays only the sample itself.  With U a unit vector or zero,
//     (1 - |e x U|) [e.U < 0]  =  1 + med3(-1, -|e x U|, -1 - e.U)
// (|e x U|^2 + (e.U)^2 = 1, so -1 - e.U >= -|e x U| exactly when e.U <= 0; U = 0 gives 0
// as the reference's GdotR < 0 test does), and CGLH = (sum W + sum W h) / sum W.  The window
// loops visit, per sample row, only the x range the tSO gate can pass, and every frame's
// load shares one 32-bit offset from a per-frame uniform base.
// ---------------------------------------------------------------------------------------
template <int FC>
__global__ void __launch_bounds__(256) k3d_rgc2(const Rgc3dArgs a)
{
    const int xM = blockIdx.x * 64 + threadIdx.x;
    const int yM = blockIdx.y * 4 + threadIdx.y;
    const int zM = blockIdx.z;
    if (xM >= a.wM || yM >= a.hM) return;
    const int w = a.w, h = a.h, nP = a.nP, M = a.M;
    const long whM = (long)a.wM * a.hM;
    const long whdM = whM * nP * a.Mz;
    const long vox = whM * zM + (long)yM * a.wM + xM;
    const long whI = 4L * w * h;
    const int wI = 2 * w;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f, s4 = 0.0f;
    if (a.count0 > 0) {
        s0 = a.out[vox];
        s1 = a.out[vox + whdM];
        s2 = a.out[vox + 2 * whdM];
        s3 = a.out[vox + 3 * whdM];
        s4 = a.prev[vox];
    }
    const float zc = (float)((zM + 0.5) / a.Mz);
    int z0;
    const float zcof = zc - 0.5f;
    if (zcof < 0.0f) z0 = 0;
    else if (zcof >= (float)(nP - 1)) z0 = nP - 2;
    else z0 = (int)floorf(zcof);
    const int zm1 = max(z0 - 1, 0), zp2 = min(z0 + 1, nP - 1);
    const float z01 = zc - ((float)z0 + 0.5f);
    const int jlo = -(int)(2.0f * a.radius), jhi = (int)(2.0f * a.radius + 1.0f);
    const int klo = -(int)a.radius, khi = (int)a.radius + 1;
    const float nFf = (float)a.nFor, nF1 = (float)(a.nFor - 1);
    // sample-point geometry (liveSRRF 3-D: xc = xM/M + 0.5/M with zero drift)
    const float xc = (float)((xM + 0.5) / M + 0.0);
    const float yc = (float)((yM + 0.5) / M + 0.0);
    const int bx = (int)(2.0f * xc), by = (int)(2.0f * yc), zi = (int)zc;
    const float tSO = a.tSO, tSO2 = tSO * tSO;
    const size_t fbytes = (size_t)nP * whI * sizeof(float4);   // one frame of U

    for (int f0 = 0; f0 < a.nframes; f0 += FC) {
        const int nf = min(FC, a.nframes - f0);
        const int c00 = a.count0 + f0;
        const char *Ub = reinterpret_cast<const char *>(a.U + (long)(a.frame0 + f0) * nP * whI);
        float acc[FC];
#pragma unroll
        for (int q = 0; q < FC; q++) acc[q] = 0.0f;
        float sumW = 0.0f;
        for (int j = jlo; j <= jhi; j++) {
            const int byj = by + j;
            const float vy = (float)byj / 2.0f;
            if (!(vy > 0.0f && vy < (float)h)) continue;
            const int m = byj - 1;                     // (int)(2 (vy - 0.5)), vy >= 0.5
            const float dy = vy - yc;
            const float dy2 = __fmul_rn(dy, dy);
            for (int k = klo; k <= khi; k++) {
                const float vz = (float)zi + (float)k;
                if (!(vz > 0.0f && vz < (float)nP)) continue;
                const float dz = vz - zc;
                const float dz2 = __fmul_rn(dz, dz);
                // x range of the tSO gate, one extra sample each side for rounding
                const float rem = tSO2 - dy2 - dz2;
                if (rem < -1e-3f * tSO2) continue;
                const float rx = sqrtf(fmaxf(rem, 0.0f));
                const int ilo = max(jlo, (int)floorf(2.0f * (xc - rx)) - bx - 1);
                const int ihi = min(jhi, (int)ceilf(2.0f * (xc + rx)) - bx + 1);
                const long rowoff = whI * (int)vz + (long)m * wI;
                for (int i = ilo; i <= ihi; i++) {
                    const int bxi = bx + i;
                    const float vx = (float)bxi / 2.0f;
                    if (!(vx > 0.0f && vx < (float)w)) continue;
                    const float dx = vx - xc;
                    float d2 = __fmul_rn(dx, dx);
                    d2 = __fadd_rn(d2, dy2);
                    d2 = __fadd_rn(d2, dz2);
                    const float d = sqrtf(d2);
                    if (!(d != 0.0f && d <= tSO)) continue;
                    float dw = d * expf(-(d * d) / a.tSS);
                    dw = dw * dw * dw * dw;
                    sumW += dw;
                    const float invd = __builtin_amdgcn_rcpf(d);
                    const float ex = dx * invd, ey = dy * invd, ez = dz * invd;
                    const unsigned off = (unsigned)(rowoff + (bxi - 1)) * (unsigned)sizeof(float4);
#pragma unroll
                    for (int q = 0; q < FC; q++) {
                        if (q < nf) {
                            const float4 u = *reinterpret_cast<const float4 *>(
                                Ub + (size_t)q * fbytes + off);
                            float c1 = fmaf(u.x, -ex, -1.0f);
                            c1 = fmaf(u.y, -ey, c1);
                            c1 = fmaf(u.z, -ez, c1);
                            const float X = fmaf(u.y, ez, -u.z * ey);
                            const float Y = fmaf(u.z, ex, -u.x * ez);
                            const float Z = fmaf(u.x, ey, -u.y * ex);
                            const float sn = __builtin_amdgcn_sqrtf(fmaf(X, X, fmaf(Y, Y, Z * Z)));
                            acc[q] = fmaf(dw, __builtin_amdgcn_fmed3f(-1.0f, -sn, c1), acc[q]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < FC; q++) {
            if (q < nf) {
                const int fi = f0 + q;
                float C = (sumW + acc[q]) / sumW;
                C = apply_sens3(C, a.sens, a.sens_int);
                const float *Lf = a.L + (long)fi * nP * whM + (long)yM * a.wM + xM;
                const float v_1 = Lf[whM * zm1], v0 = Lf[whM * z0];
                const float v1 = Lf[whM * (z0 + 1)], v2 = Lf[whM * zp2];
                // axial cubic (live3DSRRF.cl:448-452), left-to-right as written
                float aa = __fmul_rn(-0.5f, v_1);
                aa = __fadd_rn(aa, __fmul_rn(1.5f, v0));
                aa = __fsub_rn(aa, __fmul_rn(1.5f, v1));
                aa = __fadd_rn(aa, __fmul_rn(0.5f, v2));
                float bb = __fsub_rn(v_1, __fmul_rn(2.5f, v0));
                bb = __fadd_rn(bb, __fmul_rn(2.0f, v1));
                bb = __fsub_rn(bb, __fmul_rn(0.5f, v2));
                const float cc = __fadd_rn(__fmul_rn(-0.5f, v_1), __fmul_rn(0.5f, v1));
                float v = __fmul_rn(__fmul_rn(__fmul_rn(aa, z01), z01), z01);
                v = __fadd_rn(v, __fmul_rn(__fmul_rn(bb, z01), z01));
                v = __fadd_rn(v, __fmul_rn(cc, z01));
                v = __fadd_rn(v, v0);
                const float t = a.int_weighting ? v * C : C;
                const float e2 = a.int_weighting ? ((t * v) * C) : C * C;
                if (c00 + q == 0) {
                    s0 = t / nFf;
                    s1 = e2 / nFf;
                    s2 = 0.0f;
                    s3 = v / nFf;
                } else {
                    s0 = s0 + t / nFf;
                    s1 = s1 + e2 / nFf;
                    s2 = s2 + (t * s4) / nF1;
                    s3 = s3 + v / nFf;
                }
                s4 = t;   // single 1/(n-1) in TAC2 (live3DSRRF.cl:472)
            }
        }
    }
    a.out[vox] = s0;
    a.out[vox + whdM] = s1;
    a.out[vox + 2 * whdM] = s2;
    a.out[vox + 3 * whdM] = s3;
    a.prev[vox] = s4;
}

// ---------------------------------------------------------------------------------------
// Tiled 3-D RGC for frames without drift.  A block of 4 waves covers an 8x8 macro-pixel tile
// of one magnified z-slice; wave w takes lateral phase 4 blockIdx.y + w with one lane per
// macro-pixel, so every tap (LDS offset, unit direction e, weight W) is wave-uniform and is
// read from the block's LDS copy of its phases' tables.  Each frame's unit gradients over the
// planes the slice reaches are staged in LDS (rows split into even and odd columns, so a
// wave's stride-2 reads are consecutive float4s); the next frame's are loaded into registers
// while the current one is evaluated.  Per sample, with U the unit gradient (or 0):
//     c = U.e,  p = U - c e,  |e x U| = |p|,  h = med3(-1, -|p|, -1 - c)
// (h = -|e x U| when c < 0, else -1; U = 0 gives -1), and
//     CGLH = (sum W + sum W h) / sum_gated W
// equals the reference's sum over GdotR < 0 of W (1 - |e x G| / |G|) over its weight sum
// (live3DSRRF.cl:366-411).
// ---------------------------------------------------------------------------------------
constexpr int K3_FE = 16;   // tile elements each thread loads per frame (4096 per block)

__device__ __forceinline__ int axis_class3(int X0, int xs_log2)
{
    if (X0 < (1 << xs_log2)) return 0;
    return 32 - __clz(X0) - xs_log2;   // floor(log2 X0) - xs_log2 + 1
}

__device__ __forceinline__ float tap3_acc(const float4 u, const float4 q, float acc)
{
    float c = u.x * q.x;
    c = fmaf(u.y, q.y, c);
    c = fmaf(u.z, q.z, c);
    const float px = fmaf(-c, q.x, u.x);
    const float py = fmaf(-c, q.y, u.y);
    const float pz = fmaf(-c, q.z, u.z);
    float n2 = px * px;
    n2 = fmaf(py, py, n2);
    n2 = fmaf(pz, pz, n2);
    const float sn = __builtin_amdgcn_sqrtf(n2);
    return fmaf(q.w, __builtin_amdgcn_fmed3f(-1.0f, -sn, -1.0f - c), acc);
}

// SPLIT waves per phase: with 2, the two waves of a phase take the two halves of its tap
// list and meet in LDS at the frame's barrier (twice the waves per CU for the same LDS).
template <int SPLIT>
__global__ void __launch_bounds__(256 * SPLIT) k3d_rgc3(const Rgc3dArgs a, const Rgc3Tables t)
{
    constexpr int NT = 256 * SPLIT;            // threads per block
    constexpr int FE = K3_FE / SPLIT;          // tile elements each thread loads per frame
    extern __shared__ __align__(16) unsigned char smem[];
    float4 *lU = reinterpret_cast<float4 *>(smem);
    float4 *lTv = lU + t.NZ * t.PZ;
    int *lTo = reinterpret_cast<int *>(lTv + t.gtaps);
    float *lAcc = reinterpret_cast<float *>(lTo + t.gtaps);   // [4][64] second halves

    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wave = wv & 3, half = wv >> 2;   // phase slot, tap half
    const int lx = lane & 7, ly = lane >> 3;
    const int w = a.w, h = a.h, nP = a.nP, M = a.M, MM = M * M;
    const int tiles_x = (w + 7) >> 3;
    const int tX = blockIdx.x % tiles_x, tY = blockIdx.x / tiles_x;
    const int X0 = tX << 3, Y0 = tY << 3, X = X0 + lx, Y = Y0 + ly;
    const int zM = blockIdx.z;
    const int ph0 = blockIdx.y * 4;
    const int phN = min(MM, ph0 + 4);
    const int ph = ph0 + wave;
    const bool active = ph < phN;
    int cls = 0;
    if (t.ncls_x > 1 || t.ncls_y > 1)
        cls = axis_class3(Y0, t.xs_log2) * t.ncls_x + axis_class3(X0, t.xs_log2);
    const Phase3 *PB = t.phases + ((long)cls * t.nZM + zM) * MM;

    // the tap tables of the block's phases (plus the spares after them)
    const int tg0 = PB[ph0].first_tap;
    const int tg1 = PB[phN - 1].first_tap + PB[phN - 1].ntaps;
    for (int e = tid; e < tg1 - tg0 + TAP_SPARE; e += NT) {
        lTv[e] = t.tapv[tg0 + e];
        lTo[e] = t.tapo[tg0 + e];
    }

    // fill plan: tile element e -> (LDS index, offset of its unit gradient in one frame;
    // -1 outside the 0 < vx < w, 0 < vy < h gates, where the tile holds 0)
    const int zlo = PB[ph0].zlo, nz = PB[ph0].nz;
    const int wI = 2 * w, hI = 2 * h, whI = wI * hI;
    const int per_plane = t.TH * t.TW;
    const int nel = nz * per_plane;
    int gofs[FE], lidx[FE];
#pragma unroll
    for (int q = 0; q < FE; q++) {
        const int e = tid + q * NT;
        gofs[q] = -1;
        lidx[q] = -1;
        if (e < nel) {
            const int kz = e / per_plane;
            const int r = e - kz * per_plane;
            const int gy = r / t.TW, gx = r - gy * t.TW;
            const int bxi = 2 * X0 - t.G0 + gx, byj = 2 * Y0 - t.G0 + gy;
            lidx[q] = kz * t.PZ + gy * t.P + (gx & 1) * t.PW + (gx >> 1);
            if (bxi >= 1 && bxi <= wI - 1 && byj >= 1 && byj <= hI - 1)
                gofs[q] = (zlo + kz) * whI + (byj - 1) * wI + (bxi - 1);
        }
    }

    // the lane's voxel, its gated weight sum, and its cumulants
    int ft = 0, nt = 0, sx = 0, sy = 0;
    float wsum = 0.0f, wsg = 0.0f;
    if (active) {
        const Phase3 P = PB[ph];
        ft = P.first_tap - tg0;
        nt = P.ntaps;
        wsum = wsg = P.wsum;
        sx = P.sx;
        sy = P.sy;
        const bool interior = (2 * X0 + t.R2lo >= 1) && (2 * (X0 + 7) + 1 + t.R2hi <= wI - 1) &&
                              (2 * Y0 + t.R2lo >= 1) && (2 * (Y0 + 7) + 1 + t.R2hi <= hI - 1);
        if (!interior) {
            // border: only the samples inside the gates count, in the reference's order
            wsg = 0.0f;
            for (int k = 0; k < nt; k++) {
                const int kk = t.tapk[P.first_tap + k];
                const int kx = 2 * X + (kk & 255) - 128, ky = 2 * Y + ((kk >> 8) & 255) - 128;
                if (kx >= 1 && kx <= wI - 1 && ky >= 1 && ky <= hI - 1)
                    wsg += t.tapv[P.first_tap + k].w;
            }
        }
    }
    const bool inimg = active && X < w && Y < h;
    const int xM = X * M + sx, yM = Y * M + sy;
    const long whM = (long)a.wM * a.hM;
    const long whdM = whM * nP * a.Mz;
    const long vox = whM * zM + (long)yM * a.wM + xM;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f, s4 = 0.0f;
    if (a.count0 > 0 && inimg) {
        s0 = a.out[vox];
        s1 = a.out[vox + whdM];
        s2 = a.out[vox + 2 * whdM];
        s3 = a.out[vox + 3 * whdM];
        s4 = a.prev[vox];
    }
    // axial cubic planes (live3DSRRF.cl:418-452)
    const float zc = (float)((zM + 0.5) / a.Mz);
    int z0;
    const float zcof = zc - 0.5f;
    if (zcof < 0.0f) z0 = 0;
    else if (zcof >= (float)(nP - 1)) z0 = nP - 2;
    else z0 = (int)floorf(zcof);
    const int zm1 = max(z0 - 1, 0), zp2 = min(z0 + 1, nP - 1);
    const float z01 = zc - ((float)z0 + 0.5f);
    const float nFf = (float)a.nFor, nF1 = (float)(a.nFor - 1);

    const long fstride = (long)nP * whI;
    {
        const float4 *Uf = a.U + (long)a.frame0 * fstride;
#pragma unroll
        for (int q = 0; q < FE; q++)
            if (lidx[q] >= 0)
                lU[lidx[q]] = gofs[q] >= 0 ? Uf[gofs[q]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    __syncthreads();

    const int base = 2 * ly * t.P + lx;
    const float4 *Tv = lTv + ft;
    const int *To = lTo + ft;
    for (int fi = 0; fi < a.nframes; fi++) {
        const bool more = fi + 1 < a.nframes;
        // lateral intensities of this frame first, then the next frame's tile: the tile
        // loads stay in flight through the tap loop
        float v_1 = 0.0f, v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
        if (inimg && half == 0) {
            const float *Lf = a.L + (long)fi * nP * whM + (long)yM * a.wM + xM;
            v_1 = Lf[whM * zm1];
            v0 = Lf[whM * z0];
            v1 = Lf[whM * (z0 + 1)];
            v2 = Lf[whM * zp2];
        }
        float4 st[FE];
        if (more) {
            const float4 *Un = a.U + (long)(a.frame0 + fi + 1) * fstride;
#pragma unroll
            for (int q = 0; q < FE; q++)
                st[q] = gofs[q] >= 0 ? Un[gofs[q]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        float acc = 0.0f;
        if (active) {
            // this wave's share of the tap list (both halves multiples of 4)
            const int h0 = SPLIT == 1 ? 0 : (half ? ((nt >> 2) + 1) / 2 * 4 : 0);
            const int h1 = SPLIT == 1 ? nt : (half ? nt : ((nt >> 2) + 1) / 2 * 4);
#pragma unroll 1
            for (int k = h0; k < h1; k += 4) {
                const int4 o = *reinterpret_cast<const int4 *>(To + k);
                const float4 q0 = Tv[k], q1 = Tv[k + 1], q2 = Tv[k + 2], q3 = Tv[k + 3];
                const float4 u0 = lU[base + o.x], u1 = lU[base + o.y];
                const float4 u2 = lU[base + o.z], u3 = lU[base + o.w];
                acc = tap3_acc(u0, q0, acc);
                acc = tap3_acc(u1, q1, acc);
                acc = tap3_acc(u2, q2, acc);
                acc = tap3_acc(u3, q3, acc);
            }
            if (SPLIT == 2 && half) lAcc[wave * 64 + lane] = acc;
        }
        if (SPLIT == 2) __syncthreads();   // partial sums written; every wave done with the tile
        if (active && half == 0) {
            if (SPLIT == 2) acc += lAcc[wave * 64 + lane];
            if (inimg) {
                float C = (wsum + acc) / wsg;
                C = apply_sens3(C, a.sens, a.sens_int);
                // axial cubic (live3DSRRF.cl:448-452), left-to-right as written
                float aa = __fmul_rn(-0.5f, v_1);
                aa = __fadd_rn(aa, __fmul_rn(1.5f, v0));
                aa = __fsub_rn(aa, __fmul_rn(1.5f, v1));
                aa = __fadd_rn(aa, __fmul_rn(0.5f, v2));
                float bb = __fsub_rn(v_1, __fmul_rn(2.5f, v0));
                bb = __fadd_rn(bb, __fmul_rn(2.0f, v1));
                bb = __fsub_rn(bb, __fmul_rn(0.5f, v2));
                const float cc = __fadd_rn(__fmul_rn(-0.5f, v_1), __fmul_rn(0.5f, v1));
                float v = __fmul_rn(__fmul_rn(__fmul_rn(aa, z01), z01), z01);
                v = __fadd_rn(v, __fmul_rn(__fmul_rn(bb, z01), z01));
                v = __fadd_rn(v, __fmul_rn(cc, z01));
                v = __fadd_rn(v, v0);
                const float tt = a.int_weighting ? v * C : C;
                const float e2 = a.int_weighting ? ((tt * v) * C) : C * C;
                if (a.count0 + fi == 0) {
                    s0 = tt / nFf;
                    s1 = e2 / nFf;
                    s2 = 0.0f;
                    s3 = v / nFf;
                } else {
                    s0 = s0 + tt / nFf;
                    s1 = s1 + e2 / nFf;
                    s2 = s2 + (tt * s4) / nF1;
                    s3 = s3 + v / nFf;
                }
                s4 = tt;   // single 1/(n-1) in TAC2 (live3DSRRF.cl:472)
            }
        }
        if (more) {
            if (SPLIT == 1) __syncthreads();   // every wave is done with this frame's tile
#pragma unroll
            for (int q = 0; q < FE; q++)
                if (lidx[q] >= 0) lU[lidx[q]] = st[q];
            __syncthreads();
        } else if (SPLIT == 2) {
            __syncthreads();   // the partial sums are read before the next frame reuses them
        }
    }
    if (inimg && half == 0) {
        a.out[vox] = s0;
        a.out[vox + whdM] = s1;
        a.out[vox + 2 * whdM] = s2;
        a.out[vox + 3 * whdM] = s3;
        a.prev[vox] = s4;
    }
}

// One float4 of a frame's unit gradients through a bounds-checked buffer load: `off` is a
// byte offset into the frame's `nbytes`; an offset at or past the end (kOffNone) reads 0
// without touching memory, which is how the tile's gated-out elements are filled.
constexpr unsigned kOffNone = 0x80000000u;

__device__ __forceinline__ float4 load_u4_or_zero(const float4 *frame, unsigned nbytes,
                                                  unsigned off)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float4 *>(frame), (short)0, (int)nbytes, 0x00020000);
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                       __uint_as_float(v[3]));
#else
    return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#endif
}

// Tiled 3-D RGC, two z-slices per block.  Consecutive slices zM = 2k, 2k+1 reach nearly the
// same planes, so one staged tile (the union of their planes, t.NZ2) serves both: the block's
// 8 waves are (phase 4 blockIdx.y + w, slice zs), one wave per phase and slice, and each
// frame's tile is staged once for twice the samples of a one-slice block.  The staging loads
// are bounds-checked buffer loads with 32-bit offsets (a frame of gradients is below 2 GiB):
// tile elements outside the image gates read 0 through an out-of-range offset, with no
// branch or register clearing per element.  The per-voxel divisions by the weight sum and by nFrameForSRRF
// (live3DSRRF.cl:427,454-465) are multiplications by reciprocals taken once (within an ulp of
// the quotient).  The taps, their order and the sample arithmetic are those of k3d_rgc3.
template <int ZP>
__global__ void __launch_bounds__(256 * ZP) k3d_rgc3z(const Rgc3dArgs a, const Rgc3Tables t)
{
    constexpr int NT = 256 * ZP;
    constexpr int FE = K3_FE * 256 / NT;       // 4096 tile elements per block
    extern __shared__ __align__(16) unsigned char smem[];
    float4 *lU = reinterpret_cast<float4 *>(smem);
    float4 *lTv = lU + t.NZ2 * t.PZ;
    int *lTo = reinterpret_cast<int *>(lTv + ZP * t.gtaps);

    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wave = wv & 3, zs = wv >> 2;     // phase slot, slice of the pair
    const int lx = lane & 7, ly = lane >> 3;
    const int w = a.w, h = a.h, nP = a.nP, M = a.M, MM = M * M;
    const int tiles_x = (w + 7) >> 3;
    const int tX = blockIdx.x % tiles_x, tY = blockIdx.x / tiles_x;
    const int X0 = tX << 3, Y0 = tY << 3, X = X0 + lx, Y = Y0 + ly;
    const int zM0 = blockIdx.z * ZP;
    const int nzs = min(ZP, t.nZM - zM0);      // slices of this block
    const int zM = zM0 + zs;
    const int ph0 = blockIdx.y * 4;
    const int phN = min(MM, ph0 + 4);
    const int ph = ph0 + wave;
    const bool active = ph < phN && zs < nzs;
    int cls = 0;
    if (t.ncls_x > 1 || t.ncls_y > 1)
        cls = axis_class3(Y0, t.xs_log2) * t.ncls_x + axis_class3(X0, t.xs_log2);
    const Phase3 *PC = t.phases + (long)cls * t.nZM * MM;

    // the union of the slices' planes, and each slice's tap tables (phases ph0 .. phN-1)
    int zlo = 1 << 30, zhi = -1;
    for (int s = 0; s < nzs; s++) {
        const Phase3 *PB = PC + (long)(zM0 + s) * MM;
        const int tg0 = PB[ph0].first_tap;
        const int tg1 = PB[phN - 1].first_tap + PB[phN - 1].ntaps;
        for (int e = tid; e < tg1 - tg0 + TAP_SPARE; e += NT) {
            lTv[s * t.gtaps + e] = t.tapv[tg0 + e];
            lTo[s * t.gtaps + e] = t.tapo[tg0 + e];
        }
        if (PB[ph0].nz > 0) {
            zlo = min(zlo, PB[ph0].zlo);
            zhi = max(zhi, PB[ph0].zlo + PB[ph0].nz - 1);
        }
    }
    const int nzU = zhi >= zlo ? zhi - zlo + 1 : 0;

    // fill plan over the union's planes; gated-out elements are cleared here, once
    const int wI = 2 * w, hI = 2 * h, whI = wI * hI;
    const int per_plane = t.TH * t.TW;
    const int nel = nzU * per_plane;
    // goff: byte offset of the element's gradient in one frame (kOffNone outside the gates);
    // lidx: its LDS index, -1 for the thread's unused element slots
    unsigned goff[FE];
    int lidx[FE];
#pragma unroll
    for (int q = 0; q < FE; q++) {
        const int e = tid + q * NT;
        goff[q] = kOffNone;
        lidx[q] = -1;
        if (e < nel) {
            const int kz = e / per_plane;
            const int r = e - kz * per_plane;
            const int gy = r / t.TW, gx = r - gy * t.TW;
            const int bxi = 2 * X0 - t.G0 + gx, byj = 2 * Y0 - t.G0 + gy;
            lidx[q] = kz * t.PZ + gy * t.P + (gx & 1) * t.PW + (gx >> 1);
            if (bxi >= 1 && bxi <= wI - 1 && byj >= 1 && byj <= hI - 1)
                goff[q] = (unsigned)((zlo + kz) * whI + (byj - 1) * wI + (bxi - 1)) *
                          (unsigned)sizeof(float4);
        }
    }

    int ft = 0, nt = 0, sx = 0, sy = 0, zoff = 0;
    float wsum = 0.0f, wsg = 0.0f;
    if (active) {
        const Phase3 *PB = PC + (long)zM * MM;
        const Phase3 P = PB[ph];
        ft = zs * t.gtaps + P.first_tap - PB[ph0].first_tap;
        nt = P.ntaps;
        wsum = wsg = P.wsum;
        sx = P.sx;
        sy = P.sy;
        zoff = P.nz > 0 ? (P.zlo - zlo) * t.PZ : 0;
        const bool interior = (2 * X0 + t.R2lo >= 1) && (2 * (X0 + 7) + 1 + t.R2hi <= wI - 1) &&
                              (2 * Y0 + t.R2lo >= 1) && (2 * (Y0 + 7) + 1 + t.R2hi <= hI - 1);
        if (!interior) {
            wsg = 0.0f;
            for (int k = 0; k < nt; k++) {
                const int kk = t.tapk[P.first_tap + k];
                const int kx = 2 * X + (kk & 255) - 128, ky = 2 * Y + ((kk >> 8) & 255) - 128;
                if (kx >= 1 && kx <= wI - 1 && ky >= 1 && ky <= hI - 1)
                    wsg += t.tapv[P.first_tap + k].w;
            }
        }
    }
    const float rwsg = 1.0f / wsg;
    const bool inimg = active && X < w && Y < h;
    const int xM = X * M + sx, yM = Y * M + sy;
    const long whM = (long)a.wM * a.hM;
    const long whdM = whM * nP * a.Mz;
    const long vox = whM * zM + (long)yM * a.wM + xM;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f, s4 = 0.0f;
    if (a.count0 > 0 && inimg) {
        s0 = a.out[vox];
        s1 = a.out[vox + whdM];
        s2 = a.out[vox + 2 * whdM];
        s3 = a.out[vox + 3 * whdM];
        s4 = a.prev[vox];
    }
    const float zc = (float)((zM + 0.5) / a.Mz);
    int z0;
    const float zcof = zc - 0.5f;
    if (zcof < 0.0f) z0 = 0;
    else if (zcof >= (float)(nP - 1)) z0 = nP - 2;
    else z0 = (int)floorf(zcof);
    const int zm1 = max(z0 - 1, 0), zp2 = min(z0 + 1, nP - 1);
    const float z01 = zc - ((float)z0 + 0.5f);
    const float rnF = 1.0f / (float)a.nFor, rnF1 = 1.0f / (float)(a.nFor - 1);

    const long fstride = (long)nP * whI;
    const unsigned fbytes = (unsigned)(fstride * (long)sizeof(float4));
    {
        const float4 *Uf = a.U + (long)a.frame0 * fstride;
#pragma unroll
        for (int q = 0; q < FE; q++) {
            const float4 v = load_u4_or_zero(Uf, fbytes, goff[q]);
            if (lidx[q] >= 0) lU[lidx[q]] = v;
        }
    }
    __syncthreads();

    const int base = 2 * ly * t.P + lx + zoff;
    const float4 *Tv = lTv + ft;
    const int *To = lTo + ft;
    for (int fi = 0; fi < a.nframes; fi++) {
        const bool more = fi + 1 < a.nframes;
        float v_1 = 0.0f, v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
        if (inimg) {
            const float *Lf = a.L + (long)fi * nP * whM + (long)yM * a.wM + xM;
            v_1 = Lf[whM * zm1];
            v0 = Lf[whM * z0];
            v1 = Lf[whM * (z0 + 1)];
            v2 = Lf[whM * zp2];
        }
        float4 st[FE];
        if (more) {
            const float4 *Un = a.U + (long)(a.frame0 + fi + 1) * fstride;
#pragma unroll
            for (int q = 0; q < FE; q++) st[q] = load_u4_or_zero(Un, fbytes, goff[q]);
        }
        float acc = 0.0f;
        if (active) {
#pragma unroll 1
            for (int k = 0; k < nt; k += 4) {
                const int4 o = *reinterpret_cast<const int4 *>(To + k);
                const float4 q0 = Tv[k], q1 = Tv[k + 1], q2 = Tv[k + 2], q3 = Tv[k + 3];
                const float4 u0 = lU[base + o.x], u1 = lU[base + o.y];
                const float4 u2 = lU[base + o.z], u3 = lU[base + o.w];
                acc = tap3_acc(u0, q0, acc);
                acc = tap3_acc(u1, q1, acc);
                acc = tap3_acc(u2, q2, acc);
                acc = tap3_acc(u3, q3, acc);
            }
        }
        if (inimg) {
            float C = (wsum + acc) * rwsg;
            C = apply_sens3(C, a.sens, a.sens_int);
            // axial cubic (live3DSRRF.cl:448-452), left-to-right as written
            float aa = __fmul_rn(-0.5f, v_1);
            aa = __fadd_rn(aa, __fmul_rn(1.5f, v0));
            aa = __fsub_rn(aa, __fmul_rn(1.5f, v1));
            aa = __fadd_rn(aa, __fmul_rn(0.5f, v2));
            float bb = __fsub_rn(v_1, __fmul_rn(2.5f, v0));
            bb = __fadd_rn(bb, __fmul_rn(2.0f, v1));
            bb = __fsub_rn(bb, __fmul_rn(0.5f, v2));
            const float cc = __fadd_rn(__fmul_rn(-0.5f, v_1), __fmul_rn(0.5f, v1));
            float v = __fmul_rn(__fmul_rn(__fmul_rn(aa, z01), z01), z01);
            v = __fadd_rn(v, __fmul_rn(__fmul_rn(bb, z01), z01));
            v = __fadd_rn(v, __fmul_rn(cc, z01));
            v = __fadd_rn(v, v0);
            const float tt = a.int_weighting ? v * C : C;
            const float e2 = a.int_weighting ? ((tt * v) * C) : C * C;
            if (a.count0 + fi == 0) {
                s0 = tt * rnF;
                s1 = e2 * rnF;
                s2 = 0.0f;
                s3 = v * rnF;
            } else {
                s0 = s0 + tt * rnF;
                s1 = s1 + e2 * rnF;
                s2 = s2 + (tt * s4) * rnF1;
                s3 = s3 + v * rnF;
            }
            s4 = tt;   // single 1/(n-1) in TAC2 (live3DSRRF.cl:472)
        }
        if (more) {
            __syncthreads();   // every wave is done with this frame's tile
#pragma unroll
            for (int q = 0; q < FE; q++)
                if (lidx[q] >= 0) lU[lidx[q]] = st[q];
            __syncthreads();
        }
    }
    if (inimg) {
        a.out[vox] = s0;
        a.out[vox + whdM] = s1;
        a.out[vox + 2 * whdM] = s2;
        a.out[vox + 3 * whdM] = s3;
        a.prev[vox] = s4;
    }
}

size_t rgc3z_lds_bytes(const Rgc3Tables &t)
{
    return sizeof(float4) * ((size_t)t.NZ2 * t.PZ + 2 * (size_t)t.gtaps) +
           sizeof(int) * 2 * (size_t)t.gtaps;
}

size_t rgc3_lds_bytes(const Rgc3Tables &t)
{
    return sizeof(float4) * ((size_t)t.NZ * t.PZ + t.gtaps) + sizeof(int) * (size_t)t.gtaps +
           sizeof(float) * 4 * 64;
}

// Gated samples of one frame without drift, summed over all voxels.
__global__ void __launch_bounds__(256) k3d_count(int w, int h, int nP, int M, int Mz,
                                                 float radius, float tSO,
                                                 unsigned long long *total)
{
    const int wM = w * M, hM = h * M;
    const long n = (long)wM * hM * nP * Mz;
    unsigned long long cnt = 0;
    const int jlo = -(int)(2.0f * radius), jhi = (int)(2.0f * radius + 1.0f);
    const int klo = -(int)radius, khi = (int)radius + 1;
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < n;
         o += (long)gridDim.x * blockDim.x) {
        const int zM = (int)(o / ((long)wM * hM));
        const long r = o - (long)zM * wM * hM;
        const int yM = (int)(r / wM), xM = (int)(r - (long)yM * wM);
        const float xc = (float)((xM + 0.5) / M), yc = (float)((yM + 0.5) / M);
        const float zc = (float)((zM + 0.5) / Mz);
        const float bx = (float)((int)(2.0f * xc)), by = (float)((int)(2.0f * yc));
        for (int j = jlo; j <= jhi; j++) {
            const float vy = (by + (float)j) / 2.0f;
            if (!(vy > 0.0f && vy < (float)h)) continue;
            for (int i = jlo; i <= jhi; i++) {
                const float vx = (bx + (float)i) / 2.0f;
                if (!(vx > 0.0f && vx < (float)w)) continue;
                const float dx = vx - xc, dy = vy - yc;
                for (int k = klo; k <= khi; k++) {
                    const float vz = (float)((int)zc) + (float)k;
                    if (!(vz > 0.0f && vz < (float)nP)) continue;
                    const float dz = vz - zc;
                    float d2 = __fmul_rn(dx, dx);
                    d2 = __fadd_rn(d2, __fmul_rn(dy, dy));
                    d2 = __fadd_rn(d2, __fmul_rn(dz, dz));
                    const float d = sqrtf(d2);
                    if (d != 0.0f && d <= tSO) cnt++;
                }
            }
        }
    }
    __shared__ unsigned long long red[256];
    const int tid = threadIdx.x;
    red[tid] = cnt;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    if (tid == 0) atomicAdd(total, red[0]);
}

// ---------------------------------------------------------------------------------------
hipError_t launch3d_prepare(const void *src, int dtype, const Prep3dArgs &p, int n,
                            hipStream_t s)
{
    const long npx = (long)n * p.nP * p.w * p.h;
    if (npx == 0) return hipSuccess;
    hipLaunchKernelGGL(k3d_reshape, dim3(grid1(npx)), dim3(256), 0, s, src, dtype, p.planes,
                       p.W, p.H, p.w, p.h, p.nP, p.cell, p.ic, p.bg, npx);
    hipLaunchKernelGGL(k3d_align, dim3(grid1(npx)), dim3(256), 0, s, p.planes, p.aligned, p.w,
                       p.h, p.nP, p.st, npx);
    hipLaunchKernelGGL(k3d_grad, dim3(grid1(npx)), dim3(256), 0, s, p.aligned, p.g, p.w, p.h,
                       p.nP, npx);
    const long npl = (long)n * p.nP;
    hipLaunchKernelGGL(k3d_unit, dim3((2 * p.w + U3_T - 1) / U3_T, (2 * p.h + U3_T - 1) / U3_T,
                                      (unsigned)std::min(npl, 65535L)),
                       dim3(256), 0, s, p.g, p.U, p.w, p.h, npl);
    return hipGetLastError();
}

hipError_t launch3d_reshape(const void *src, int dtype, const Prep3dArgs &p, int n,
                            hipStream_t s)
{
    const long npx = (long)n * p.nP * p.w * p.h;
    if (npx == 0) return hipSuccess;
    hipLaunchKernelGGL(k3d_reshape, dim3(grid1(npx)), dim3(256), 0, s, src, dtype, p.planes,
                       p.W, p.H, p.w, p.h, p.nP, p.cell, p.ic, p.bg, npx);
    return hipGetLastError();
}

hipError_t launch3d_gint(const float *g, float *out, int w, int h, int nP, long gn,
                         hipStream_t s)
{
    const long n = (long)nP * 3 * 4 * w * h;
    hipLaunchKernelGGL(k3d_gint, dim3(grid1(n)), dim3(256), 0, s, g, out, w, h, nP, gn);
    return hipGetLastError();
}

hipError_t launch3d_lateral(const Rgc3dArgs &a, hipStream_t s)
{
    if (a.nframes <= 0) return hipSuccess;
    const int nfz = a.nframes * a.nP;
    hipLaunchKernelGGL(k3d_lateral, dim3((a.wM + LAT_X - 1) / LAT_X, (a.hM + LAT_Y - 1) / LAT_Y,
                                         std::min(nfz, 65535)),
                       dim3(256), 0, s, a.px, a.L, a.w, a.h, a.nP, a.M, a.frame0, a.nframes,
                       a.drift, a.count0, a.nFor);
    return hipGetLastError();
}

hipError_t launch3d_rgc(const Rgc3dArgs &a, bool drift_zero, const Rgc3Tables *t,
                        size_t lds_limit, hipStream_t s, const char **kname, int kern,
                        int split, int zpair)
{
    if (a.nframes <= 0) return hipSuccess;
    const dim3 grid((a.wM + 63) / 64, (a.hM + 3) / 4, a.nP * a.Mz);
    // (the fast kernel addresses one frame of U with 32-bit byte offsets)
    const bool off32 = (double)a.nP * 4.0 * a.w * a.h * sizeof(float4) < 4294967296.0;
    // the tiled kernel addresses one frame of U with 32-bit element offsets
    const bool off31 = (double)a.nP * 4.0 * a.w * a.h < 2147483648.0;
    const int tiles = ((a.w + 7) >> 3) * ((a.h + 7) >> 3);
    // the two-slice kernel's buffer loads address one frame of U with 31-bit byte offsets
    const bool buf31 = (double)a.nP * 4.0 * a.w * a.h * sizeof(float4) < 2147483648.0;
    if (drift_zero && t && kern == 3 && buf31 && zpair && t->NZ2 > 0 &&
        rgc3z_lds_bytes(*t) <= lds_limit) {
        const int nz2 = (a.nP * a.Mz + 1) / 2;
        const dim3 g3(tiles, (a.M * a.M + 3) / 4, nz2);
        hipLaunchKernelGGL(k3d_rgc3z<2>, g3, dim3(512), rgc3z_lds_bytes(*t), s, a, *t);
        if (kname) *kname = "k3d_rgc3z";
    } else if (drift_zero && t && kern == 3 && off31 && rgc3_lds_bytes(*t) <= lds_limit) {
        const dim3 g3(tiles, (a.M * a.M + 3) / 4, a.nP * a.Mz);
        if (split == 1)
            hipLaunchKernelGGL(k3d_rgc3<1>, g3, dim3(256), rgc3_lds_bytes(*t), s, a, *t);
        else
            hipLaunchKernelGGL(k3d_rgc3<2>, g3, dim3(512), rgc3_lds_bytes(*t), s, a, *t);
        if (kname) *kname = "k3d_rgc3";
    } else if (drift_zero && off32 && kern >= 2) {
        hipLaunchKernelGGL(k3d_rgc2<8>, grid, dim3(64, 4), 0, s, a);
        if (kname) *kname = "k3d_rgc2";
    } else if (drift_zero) {
        hipLaunchKernelGGL(k3d_rgc<8>, grid, dim3(64, 4), 0, s, a);
        if (kname) *kname = "k3d_rgc";
    } else {
        hipLaunchKernelGGL(k3d_rgc<1>, grid, dim3(64, 4), 0, s, a);
        if (kname) *kname = "k3d_rgc";
    }
    return hipGetLastError();
}

hipError_t launch3d_count(int w, int h, int nP, int M, int Mz, float radius, float tSO,
                          unsigned long long *d_total, hipStream_t s)
{
    const long n = (long)w * M * h * M * nP * Mz;
    hipLaunchKernelGGL(k3d_count, dim3(grid1(n)), dim3(256), 0, s, w, h, nP, M, Mz, radius,
                       tSO, d_total);
    return hipGetLastError();
}

}  // namespace esrrf
