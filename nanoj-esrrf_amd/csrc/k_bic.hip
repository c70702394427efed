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
#include "esrrf_internal.h"

namespace esrrf {

// Catmull-Rom kernel, same expression order as liveSRRF.cl:26-35 (exact for the half-grid
// arguments the gradient interpolation uses).
__device__ __forceinline__ float cubic_w(float x)
{
    const float a = 0.5f;
    if (x < 0.0f) x = -x;
    float z = 0.0f;
    if (x < 1.0f) {
        float t = __fmul_rn(x, -a + 2.0f);
        t = __fadd_rn(t, a - 3.0f);
        z = __fmul_rn(x, x);
        z = __fmul_rn(z, t);
        z = __fadd_rn(z, 1.0f);
    } else if (x < 2.0f) {
        float t1 = __fmul_rn(__fmul_rn(__fmul_rn(-a, x), x), x);
        float t2 = __fmul_rn(__fmul_rn(5.0f * a, x), x);
        float t3 = __fmul_rn(8.0f * a, x);
        z = __fadd_rn(__fsub_rn(__fadd_rn(t1, t2), t3), 4.0f * a);
    }
    return z;
}

// getInterpolatedValue on a plane stored in a (possibly offset) tile:
// value of image pixel (u, v) is tile[(v - oy) * pitch + (u - ox)].
__device__ __forceinline__ float interp_tile(const float *t, int pitch, int ox, int oy, int w,
                                             int h, float x, float y)
{
    const int u0 = (int)floorf(x);
    const int v0 = (int)floorf(y);
    float q = 0.0f;
    if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2) {
#pragma unroll
        for (int j = 0; j <= 3; j++) {
            const int v = v0 - 1 + j;
            float p = 0.0f;
#pragma unroll
            for (int i = 0; i <= 3; i++) {
                const int u = u0 - 1 + i;
                p = p + t[(v - oy) * pitch + (u - ox)] * cubic_w(x - (float)u);
            }
            q = q + p * cubic_w(y - (float)v);
        }
    } else {
        const int xbase = (int)fminf((float)w - 2.0f, fmaxf(x, 0.0f));
        const int ybase = (int)fminf((float)h - 2.0f, fmaxf(y, 0.0f));
        const float xf = x - (float)xbase;
        const float yf = y - (float)ybase;
        const float ll = t[(ybase - oy) * pitch + (xbase - ox)];
        const float lr = t[(ybase - oy) * pitch + (xbase + 1 - ox)];
        const float ur = t[(ybase + 1 - oy) * pitch + (xbase + 1 - ox)];
        const float ul = t[(ybase + 1 - oy) * pitch + (xbase - ox)];
        const float ua = ul + xf * (ur - ul);
        const float la = ll + xf * (lr - ll);
        q = la + yf * (ua - la);
    }
    return q;
}

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

__device__ __forceinline__ float2 unit_or_zero(float tx, float ty)
{
    const float m = fmaxf(fabsf(tx), fabsf(ty));
    if (!(m > 0.0f) || !(m <= 3.4e38f)) return make_float2(0.0f, 0.0f);
    const float sx = tx / m, sy = ty / m;
    const float r = rsqrtf(sx * sx + sy * sy);
    return make_float2(sx * r, sy * r);
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
        const float tx = interp_tile(lgx, PREP_G, gx0, gy0, w, h, (float)kx / 2.0f,
                                     (float)(ky - 1) / 2.0f);
        const float ty = interp_tile(lgy, PREP_G, gx0, gy0, w, h, (float)(kx - 1) / 2.0f,
                                     (float)ky / 2.0f);
        u = unit_or_zero(tx, ty);
    }
    U[(long)f * wI * hI + (long)ky * wI + kx] = u;
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
    const float fx = (float)x / 2.0f, fy = (float)y / 2.0f;
    gxi[(long)y * 2 * w + x] = interp_tile(lgx, PREP_G, gx0, gy0, w, h, fx, fy);
    gyi[(long)y * 2 * w + x] = interp_tile(lgy, PREP_G, gx0, gy0, w, h, fx, fy);
}

hipError_t launch_bic_prep(const float *px, float2 *U, int w, int h, int nF, hipStream_t s)
{
    if (nF <= 0) return hipSuccess;
    dim3 grid((2 * w + PREP_T - 1) / PREP_T, (2 * h + PREP_T - 1) / PREP_T, nF);
    hipLaunchKernelGGL(k_bic_prep, grid, dim3(PREP_T, PREP_T), 0, s, px, U, w, h);
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
constexpr int RAW_P = 13;

__device__ __forceinline__ int axis_class(int X0, int xs_log2)
{
    if (X0 < (1 << xs_log2)) return 0;
    return 32 - __clz(X0) - xs_log2;   // floor(log2 X0) - xs_log2 + 1
}

__device__ __forceinline__ float apply_sens(float C, float sens, int sens_int)
{
    if (!(C >= 0.0f)) return 0.0f;   // negative or NaN -> 0 (liveSRRF.cl:358-359)
    switch (sens_int) {
    case 1: return C;
    case 2: return C * C;
    case 3: return C * C * C;
    case 4: { const float c2 = C * C; return c2 * c2; }
    default: return powf(C, sens);
    }
}

// One RGC sample: acc += W * med3(-1, -|s|, -c - 1), i.e. W * ((1 - |sin|)[cos < 0] - 1)
// with c = U.(a,b) and s = U x (a,b) (liveSRRF.cl:296-341 with G normalised).
__device__ __forceinline__ float tap_acc(float2 u, const Tap &tp, float acc)
{
    float c2 = fmaf(u.x, -tp.a, -1.0f);
    c2 = fmaf(u.y, -tp.b, c2);
    float s = u.y * tp.a;
    s = fmaf(u.x, -tp.b, s);
    const float hh = __builtin_amdgcn_fmed3f(-1.0f, -fabsf(s), c2);
    return fmaf(tp.W, hh, acc);
}

template <int PPW>
__global__ void __launch_bounds__(PPW <= 4 ? 1024 : 512)
__attribute__((amdgpu_waves_per_eu(PPW <= 4 ? 4 : 2))) k_bic_rgc(const BicRgcArgs a,
                                                 const Tap *__restrict__ taps,
                                                 const Phase *__restrict__ phases,
                                                 const int *__restrict__ tapk)
{
    extern __shared__ __align__(16) unsigned char smem[];
    float2 *lU = reinterpret_cast<float2 *>(smem);
    const int usz = 4 * a.PP * a.PH;
    float *lR = reinterpret_cast<float *>(lU + usz);

    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lx = lane & 7, ly = lane >> 3;
    const int tX = blockIdx.x % a.tiles_x, tY = blockIdx.x / a.tiles_x;
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

    float st[PPW][5];
    float invw[PPW];
    int ibase[PPW];   // LDS index of the lane's bicubic support, -1: bilinear border path
#pragma unroll
    for (int p = 0; p < PPW; p++) {
        const int ph = wave + p * a.nw;
        invw[p] = 0.0f;
        ibase[p] = -1;
        st[p][0] = st[p][1] = st[p][2] = st[p][3] = st[p][4] = 0.0f;
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
            const int u0 = X + P.du, v0 = Y + P.dv;
            if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2)
                ibase[p] = (v0 - 1 - (Y0 - 2)) * RAW_P + (u0 - 1 - (X0 - 2));
            if (!a.first_init && inimg) {
                const long pix = (long)(Y * M + P.sy) * a.wM + (X * M + P.sx);
                st[p][0] = a.out[pix];
                st[p][1] = a.out[pix + whM];
                st[p][2] = a.out[pix + 2 * whM];
                st[p][3] = a.out[pix + 3 * whM];
                st[p][4] = a.prev[pix];
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

    for (int fi = 0; fi < a.nframes; fi++) {
        const int f = a.frame0 + fi;
        __syncthreads();
        {
            const float2 *Uf = a.U + whI * f;
            const int n = a.TW * a.TH;
            for (int e = tid; e < n; e += nthr) {
                const int gy = e / a.TW, gx = e - gy * a.TW;
                const int kx = kxo + gx, ky = kyo + gy;
                float2 v = make_float2(0.0f, 0.0f);
                if (kx >= 0 && kx < wI && ky >= 0 && ky < hI) v = Uf[(long)ky * wI + kx];
                lU[((gy & 1) * 2 + (gx & 1)) * planesz + (gy >> 1) * a.PP + (gx >> 1)] = v;
            }
            const float *Pf = a.px + wh * f;
            for (int e = tid; e < RAW_T * RAW_T; e += nthr) {
                const int r = e / RAW_T, c = e - r * RAW_T;
                const int ix = min(max(X0 - 2 + c, 0), w - 1);
                const int iy = min(max(Y0 - 2 + r, 0), h - 1);
                lR[r * RAW_P + c] = Pf[(long)iy * w + ix];
            }
        }
        __syncthreads();
        const bool first = (fi == 0) && a.first_init;
        // Re-read the per-phase tables every frame and hide the lane coordinates from
        // loop-invariant code motion: hoisting PPW copies of the phase geometry out of the
        // frame loop costs far more registers than re-reading it through the scalar cache.
        // (a.zero is 0 at run time; the compiler cannot prove it, so nothing below moves
        // out of the frame loop)
        const int zf = a.zero * fi;
        const int Xl = X + zf, Yl = Y + zf;
        const Phase *PHf = PHS + zf;
        const Tap *TAPf = taps + zf;

#pragma unroll
        for (int p = 0; p < PPW; p++) {
            const int ph = wave + p * a.nw;
            if (ph < MM) {
                const Phase P = PHf[ph];
                const Tap *T = TAPf + P.first_tap;
                // Tap lists are padded to a multiple of 4 with W = 0 taps.  The constants
                // of chunk c+1 are fetched (scalar cache) while the LDS reads of chunk c are
                // in flight, so one wait covers both latencies.
                float acc = 0.0f;
                Tap c0 = T[0], c1 = T[1], c2 = T[2], c3 = T[3];
                const int nch = P.ntaps >> 2;
                for (int c = 0; c < nch; c++) {
                    const float2 u0 = lU[base + c0.off];
                    const float2 u1 = lU[base + c1.off];
                    const float2 u2 = lU[base + c2.off];
                    const float2 u3 = lU[base + c3.off];
                    const int cn = (c + 1 < nch) ? 4 * (c + 1) : 4 * c;
                    const Tap n0 = T[cn], n1 = T[cn + 1], n2 = T[cn + 2], n3 = T[cn + 3];
                    acc = tap_acc(u0, c0, acc);
                    acc = tap_acc(u1, c1, acc);
                    acc = tap_acc(u2, c2, acc);
                    acc = tap_acc(u3, c3, acc);
                    c0 = n0;
                    c1 = n1;
                    c2 = n2;
                    c3 = n3;
                }
                float C = (P.wsum + acc) * invw[p];
                C = apply_sens(C, a.sens, a.sens_int);

                // interpolated intensity at (xM/M - 0.5, yM/M - 0.5) (liveSRRF.cl:361)
                float v;
                if (ibase[p] >= 0) {
                    const float *row = lR + ibase[p];
                    float q = 0.0f;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        float pp = 0.0f;
                        pp = pp + row[j * RAW_P + 0] * P.wx[0];
                        pp = pp + row[j * RAW_P + 1] * P.wx[1];
                        pp = pp + row[j * RAW_P + 2] * P.wx[2];
                        pp = pp + row[j * RAW_P + 3] * P.wx[3];
                        q = q + pp * P.wy[j];
                    }
                    v = q;
                } else {
                    const float xf = (float)(Xl * M + P.sx) / (float)M - 0.5f;
                    const float yf = (float)(Yl * M + P.sy) / (float)M - 0.5f;
                    v = interp_tile(lR, RAW_P, X0 - 2, Y0 - 2, w, h, xf, yf);
                }

                // temporal accumulation (liveSRRF.cl:363-393)
                const float t = intw ? v * C : C;
                const float e2 = intw ? (t * v) * C : C * C;
                if (first) {
                    st[p][0] = t * rn;
                    st[p][1] = e2 * rn;
                    st[p][2] = 0.0f;
                    st[p][3] = v * rn;
                } else {
                    st[p][0] = st[p][0] + t * rn;
                    st[p][1] = st[p][1] + e2 * rn;
                    st[p][2] = st[p][2] + (t * st[p][4]) * rn1;
                    st[p][3] = st[p][3] + v * rn;
                }
                st[p][4] = prev_scaled ? t * rn1 : t;
            }
        }
    }

    if (inimg) {
#pragma unroll
        for (int p = 0; p < PPW; p++) {
            const int ph = wave + p * a.nw;
            if (ph >= MM) continue;
            const Phase P = PHS[ph];
            const long pix = (long)(Y * M + P.sy) * a.wM + (X * M + P.sx);
            a.out[pix] = st[p][0];
            a.out[pix + whM] = st[p][1];
            a.out[pix + 2 * whM] = st[p][2];
            a.out[pix + 3 * whM] = st[p][3];
            a.prev[pix] = st[p][4];
        }
    }
}

template <int PPW>
static hipError_t launch_ppw(const BicRgcArgs &a, hipStream_t s)
{
    const size_t lds = sizeof(float2) * 4 * a.PP * a.PH + sizeof(float) * RAW_T * RAW_P;
    hipLaunchKernelGGL(k_bic_rgc<PPW>, dim3(a.tiles_x * a.tiles_y), dim3(64 * a.nw), lds, s, a,
                       a.taps, a.phases, a.tapk);
    return hipGetLastError();
}

hipError_t launch_bic_rgc(const BicRgcArgs &a, int ppw, hipStream_t s)
{
    if (a.nframes <= 0) return hipSuccess;
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
__device__ __forceinline__ float interp_global(const float *__restrict__ a, int w, int h,
                                               float x, float y)
{
    const int u0 = (int)floorf(x);
    const int v0 = (int)floorf(y);
    float q = 0.0f;
    if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2) {
        for (int j = 0; j <= 3; j++) {
            const int v = v0 - 1 + j;
            float p = 0.0f;
            for (int i = 0; i <= 3; i++) {
                const int u = u0 - 1 + i;
                p = p + a[(long)v * w + u] * cubic_w(x - (float)u);
            }
            q = q + p * cubic_w(y - (float)v);
        }
    } else {
        const int xbase = (int)fminf((float)w - 2.0f, fmaxf(x, 0.0f));
        const int ybase = (int)fminf((float)h - 2.0f, fmaxf(y, 0.0f));
        const float xf = x - (float)xbase;
        const float yf = y - (float)ybase;
        const float ll = a[(long)ybase * w + xbase];
        const float lr = a[(long)ybase * w + xbase + 1];
        const float ur = a[(long)(ybase + 1) * w + xbase + 1];
        const float ul = a[(long)(ybase + 1) * w + xbase];
        const float ua = ul + xf * (ur - ul);
        const float la = ll + xf * (lr - ll);
        q = la + yf * (ua - la);
    }
    return q;
}

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
