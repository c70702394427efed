// k_fht.hip -- the Fourier ("FHT") mode: Hartley-domain upsampling of each frame plus the
// rotated-Roberts-cross gradient and the radial gradient convergence in magnified space
// (resources/liveSRRF_noInterpolation.cl, FourierInterpolation.java).
//
// Upsampling of one N x N frame (N a power of two after mirror padding) by M:
//   H = Re F - Im F, F = FFT2_N(frame)                    (forward 2-D Hartley transform)
//   H' = H zero-inserted into an MN x MN spectrum, Nyquist on the negative side
//   u  = (Re - Im) FFT2_MN(H') / (MN)^2                   (normalised inverse Hartley)
// The MN transform of the sparse H' is never formed: writing x = M a + b, y = M c + d,
//   FFT2_MN(H')[Mc+d][Ma+b] = FFT2_N( H(ky,kx) e^{-2 pi i (k(ky) d + k(kx) b)/(MN)} )[c][a]
// with k() the signed frequency, so one frame costs 1 + M^2 batched N x N FFTs, all power
// of two, done by a radix-2 Stockham kernel in LDS (G signals of one batch per block).
#include "device_interp.h"
#include "esrrf_internal.h"
#include "fht_internal.h"

#include <algorithm>

namespace esrrf {

// ---------------------------------------------------------------------------------------
// Mirror padding to the N x N square (FourierInterpolation.java:118-172); the identity
// for power-of-two squares.
// ---------------------------------------------------------------------------------------
__global__ void k_fht_pad(const float *__restrict__ src, float2 *__restrict__ dst, int w, int h,
                          int N, int scaleW, int scaleH, int xROI, int yROI, long src_stride,
                          long dst_stride)
{
    const int f = blockIdx.y;
    const long n = (long)N * N;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x) {
        const int yy = (int)(e / N), xx = (int)(e - (long)yy * N);
        const int X = xx + xROI, Y = yy + yROI;
        const int sW = X / w, i = X - sW * w;
        const int sH = Y / h, j = Y - sH * h;
        const int p = ((sW - (scaleW - 1) / 2) % 2 == 0) ? i : w - 1 - i;
        const int q = ((sH - (scaleH - 1) / 2) % 2 == 0) ? j : h - 1 - j;
        dst[dst_stride * f + e] = make_float2(src[src_stride * f + (long)q * w + p], 0.0f);
    }
}

// ---------------------------------------------------------------------------------------
// Batched radix-2 Stockham FFT (forward, e^{-i}).  Signal g of batch b, element k is at
// base[b*bstride + g*gstride + k*kstride].  A block handles G consecutive signals.
// ---------------------------------------------------------------------------------------
template <int G, bool ROWS>
__global__ void __launch_bounds__(256) k_fft(float2 *__restrict__ base, int N, int log2N,
                                             long gstride, long kstride, long bstride,
                                             const float2 *__restrict__ tw)
{
    extern __shared__ float2 sm[];
    float2 *A = sm, *B = sm + G * N;
    const int g0 = blockIdx.x * G;
    float2 *data = base + bstride * blockIdx.y;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int total = G * N;
    // coalesced staging: along a row consecutive threads take consecutive elements, along
    // columns consecutive signals
    for (int e = tid; e < total; e += nthr) {
        const int g = ROWS ? e / N : e % G, k = ROWS ? e % N : e / G;
        A[g * N + k] = data[(long)(g0 + g) * gstride + (long)k * kstride];
    }
    __syncthreads();
    const int half = N >> 1;
    for (int L = 1, s = 0; s < log2N; L <<= 1, s++) {
        const int twstep = half / L;   // e^{-2 pi i k/(2L)} = tw[k * N/(2L)]
        for (int e = tid; e < G * half; e += nthr) {
            const int g = e / half, j = e - g * half;
            const int k = j & (L - 1);
            const float2 wv = tw[k * twstep];
            const float2 u = A[g * N + j];
            const float2 x = A[g * N + j + half];
            const float2 v = make_float2(x.x * wv.x - x.y * wv.y, x.x * wv.y + x.y * wv.x);
            const int o = g * N + 2 * j - k;
            B[o] = make_float2(u.x + v.x, u.y + v.y);
            B[o + L] = make_float2(u.x - v.x, u.y - v.y);
        }
        __syncthreads();
        float2 *t = A;
        A = B;
        B = t;
    }
    for (int e = tid; e < total; e += nthr) {
        const int g = ROWS ? e / N : e % G, k = ROWS ? e % N : e / G;
        data[(long)(g0 + g) * gstride + (long)k * kstride] = A[g * N + k];
    }
}

// Forward spectrum -> Hartley transform H = Re F - Im F, then the M^2 twiddled copies
// Y_bd(ky, kx) = H e^{-2 pi i (k(kx) b + k(ky) d)/(MN)} (batch index = d*M + b).
__global__ void k_fht_twiddle(const float2 *__restrict__ F, float2 *__restrict__ Y, int N, int M)
{
    const long n = (long)N * N;
    const int ph = blockIdx.y;
    const int d = ph / M, b = ph - d * M;
    const double scale = -2.0 * 3.14159265358979323846 / ((double)M * N);
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x) {
        const int ky = (int)(e / N), kx = (int)(e - (long)ky * N);
        const int sx = kx < N / 2 ? kx : kx - N;
        const int sy = ky < N / 2 ? ky : ky - N;
        const float2 f = F[e];
        const float hv = f.x - f.y;
        // phase reduced modulo MN in integers, then one sincos in double for accuracy
        const long MN = (long)M * N;
        long ip = ((long)sx * b + (long)sy * d) % MN;
        if (ip < 0) ip += MN;
        double s, c;
        sincos(scale * (double)ip, &s, &c);
        Y[(long)ph * n + e] = make_float2((float)(hv * c), (float)(hv * s));
    }
}

// (Re - Im) / (MN)^2 of phase (b, d) at (a, c) -> magnified pixel (M a + b - xROI,
// M c + d - yROI) of the cropped frame.
__global__ void k_fht_scatter(const float2 *__restrict__ Z, float *__restrict__ out, int N,
                              int M, int wM, int hM, int xROI, int yROI, float inv)
{
    const long n = (long)N * N;
    const int ph = blockIdx.y;
    const int d = ph / M, b = ph - d * M;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long)gridDim.x * blockDim.x) {
        const int c = (int)(e / N), a = (int)(e - (long)c * N);
        const int X = M * a + b - xROI, Y = M * c + d - yROI;
        if (X < 0 || X >= wM || Y < 0 || Y >= hM) continue;
        const float2 z = Z[(long)ph * n + e];
        out[(long)Y * wM + X] = (z.x - z.y) * inv;
    }
}

static int grid_of(long n)
{
    long g = (n + 255) / 256;
    return (int)(g > 4096 ? 4096 : (g < 1 ? 1 : g));
}

template <bool ROWS>
static void fft_pass(float2 *data, int N, int log2N, int batches, const float2 *tw, hipStream_t s)
{
    const long n = (long)N * N;
    const long gs = ROWS ? N : 1, ks = ROWS ? 1 : N;
    // G signals per block (G <= N so the grid N / G is whole) with 2*G*N complex values of
    // LDS <= 64 KiB
    const int G = std::max(1, std::min(std::min(16, N), 4096 / N));
    const size_t lds = 2 * sizeof(float2) * (size_t)G * N;
    const dim3 grid(N / G, batches);
    switch (G) {
    case 16: hipLaunchKernelGGL((k_fft<16, ROWS>), grid, dim3(256), lds, s, data, N, log2N, gs, ks, n, tw); break;
    case 8: hipLaunchKernelGGL((k_fft<8, ROWS>), grid, dim3(256), lds, s, data, N, log2N, gs, ks, n, tw); break;
    case 4: hipLaunchKernelGGL((k_fft<4, ROWS>), grid, dim3(256), lds, s, data, N, log2N, gs, ks, n, tw); break;
    case 2: hipLaunchKernelGGL((k_fft<2, ROWS>), grid, dim3(256), lds, s, data, N, log2N, gs, ks, n, tw); break;
    default: hipLaunchKernelGGL((k_fft<1, ROWS>), grid, dim3(256), lds, s, data, N, log2N, gs, ks, n, tw); break;
    }
}

// 2-D FFT of `batches` consecutive N x N complex arrays (N a power of two, 2..4096).
hipError_t fft2_batched(float2 *data, int N, int log2N, int batches, const float2 *tw,
                        hipStream_t s)
{
    fft_pass<true>(data, N, log2N, batches, tw, s);
    fft_pass<false>(data, N, log2N, batches, tw, s);
    return hipGetLastError();
}

hipError_t launch_fht_upsample(const FhtPlan &p, const float *frame, float *out, hipStream_t s)
{
    const long n = (long)p.N * p.N;
    hipLaunchKernelGGL(k_fht_pad, dim3(grid_of(n), 1), dim3(256), 0, s, frame, p.F, p.w, p.h,
                       p.N, p.scaleW, p.scaleH, p.xROIp, p.yROIp, 0L, 0L);
    hipError_t st = fft2_batched(p.F, p.N, p.log2N, 1, p.tw, s);
    if (st != hipSuccess) return st;
    const int MM = p.M * p.M;
    hipLaunchKernelGGL(k_fht_twiddle, dim3(grid_of(n), MM), dim3(256), 0, s, p.F, p.Y, p.N, p.M);
    st = fft2_batched(p.Y, p.N, p.log2N, MM, p.tw, s);
    if (st != hipSuccess) return st;
    const float inv = (float)(1.0 / ((double)p.M * p.N * (double)p.M * p.N));
    hipLaunchKernelGGL(k_fht_scatter, dim3(grid_of(n), MM), dim3(256), 0, s, p.Y, out, p.N, p.M,
                       p.M * p.w, p.M * p.h, p.xROI, p.yROI, inv);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Rotated Roberts cross (liveSRRF_noInterpolation.cl:138-159) and the unit tap gradients
//   T(n, m) = (Gx[m][min(n+1, W-1)], Gy[min(m+1, H-1)][n])     (:291-292, clamped reads)
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void robx(const float *P, int W, int H, int x1, int y1, float &gx,
                                     float &gy)
{
    const int x0 = x1 - 1 < 0 ? 0 : x1 - 1;
    const int y0 = y1 - 1 < 0 ? 0 : y1 - 1;
    const float p01 = P[(long)y0 * W + x1], p10 = P[(long)y1 * W + x0];
    const float p11 = P[(long)y1 * W + x1], p00 = P[(long)y0 * W + x0];
    gx = ((p01 - p10) + p11) - p00;
    gy = ((-p01 + p10) + p11) - p00;
}

__device__ __forceinline__ float2 unit_or_zero_f(float tx, float ty)
{
    const float m = fmaxf(fabsf(tx), fabsf(ty));
    if (!(m > 0.0f) || !(m <= 3.4e38f)) return make_float2(0.0f, 0.0f);
    const float sx = tx / m, sy = ty / m;
    const float r = rsqrtf(sx * sx + sy * sy);
    return make_float2(sx * r, sy * r);
}

__global__ void k_fht_prep(const float *__restrict__ px, float2 *__restrict__ U, int W, int H)
{
    const int f = blockIdx.y;
    const long wh = (long)W * H;
    const float *P = px + wh * f;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < wh;
         e += (long)gridDim.x * blockDim.x) {
        const int m = (int)(e / W), n = (int)(e - (long)m * W);
        float gx, gy, dummy;
        robx(P, W, H, min(n + 1, W - 1), m, gx, dummy);
        robx(P, W, H, n, min(m + 1, H - 1), dummy, gy);
        U[wh * f + e] = unit_or_zero_f(gx, gy);
    }
}

__global__ void k_fht_grad(const float *__restrict__ px, float *__restrict__ gx,
                           float *__restrict__ gy, int W, int H)
{
    const long wh = (long)W * H;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < wh;
         e += (long)gridDim.x * blockDim.x) {
        const int y = (int)(e / W), x = (int)(e - (long)y * W);
        float a, b;
        robx(px, W, H, x, y, a, b);
        gx[e] = a;
        gy[e] = b;
    }
}

hipError_t launch_fht_prep(const float *px, float2 *U, int W, int H, int nF, hipStream_t s)
{
    hipLaunchKernelGGL(k_fht_prep, dim3(grid_of((long)W * H), nF), dim3(256), 0, s, px, U, W, H);
    return hipGetLastError();
}

hipError_t launch_fht_grad(const float *px, float *gx, float *gy, int W, int H, hipStream_t s)
{
    hipLaunchKernelGGL(k_fht_grad, dim3(grid_of((long)W * H)), dim3(256), 0, s, px, gx, gy, W, H);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Tiled RGC in magnified space (liveSRRF_noInterpolation.cl:245-393), unshifted frames:
// the sample offsets are the integers (i, j) of the window, identical for every pixel, so
// one tap table serves the whole image.  Block = 16 x 16 pixels (4 waves of 8 x 8), the
// frame's U tile (16 + window) staged in LDS with a pitch of 8 (mod 32) float2, which
// keeps the 8 x 8 lanes of a read on distinct banks.
// ---------------------------------------------------------------------------------------
constexpr int FT = 16;

__device__ __forceinline__ float apply_sens_f(float C, float sens, int sens_int)
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

__device__ __forceinline__ float tap_acc_f(float2 u, const Tap &tp, float acc)
{
    float c2 = fmaf(u.x, -tp.a, -1.0f);
    c2 = fmaf(u.y, -tp.b, c2);
    float s = u.y * tp.a;
    s = fmaf(u.x, -tp.b, s);
    const float hh = __builtin_amdgcn_fmed3f(-1.0f, -fabsf(s), c2);
    return fmaf(tp.W, hh, acc);
}

__global__ void __launch_bounds__(256) k_fht_rgc(const FhtRgcArgs a,
                                                 const Tap *__restrict__ taps,
                                                 const int *__restrict__ tapk)
{
    extern __shared__ __align__(16) unsigned char smem[];
    float2 *lU = reinterpret_cast<float2 *>(smem);
    Tap *lT = reinterpret_cast<Tap *>(lU + a.TH * a.pitch);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lx = (lane & 7) + 8 * (wave & 1), ly = (lane >> 3) + 8 * (wave >> 1);
    const int tX = blockIdx.x % a.tiles_x, tY = blockIdx.x / a.tiles_x;
    const int X0 = tX * FT, Y0 = tY * FT;
    const int xM = X0 + lx, yM = Y0 + ly;
    const int W = a.W, H = a.H;
    const bool inimg = xM < W && yM < H;
    const long wh = (long)W * H;
    for (int e = tid; e < a.ntaps + 4; e += 256) lT[e] = taps[e];
    const bool interior = X0 - a.R >= 0 && X0 + FT - 1 + a.R + 1 <= W - 1 && Y0 - a.R >= 0 &&
                          Y0 + FT - 1 + a.R + 1 <= H - 1;
    float ws = a.wsum;
    if (!interior) {
        ws = 0.0f;
        for (int t = 0; t < a.ntaps; t++) {
            const int k = tapk[t];
            const int n = xM + (k & 255) - 128, m = yM + ((k >> 8) & 255) - 128;
            if (n >= 0 && n <= W - 1 && m >= 0 && m <= H - 1) ws += taps[t].W;
        }
    }
    const float invw = 1.0f / ws;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f, s4 = 0.0f;
    const long pix = (long)yM * W + xM;
    if (!a.first_init && inimg) {
        s0 = a.out[pix];
        s1 = a.out[pix + wh];
        s2 = a.out[pix + 2 * wh];
        s3 = a.out[pix + 3 * wh];
        s4 = a.prev[pix];
    }
    const int base = (ly + a.R) * a.pitch + (lx + a.R);
    const int nU = a.TH * a.TW;
    for (int fi = 0; fi < a.nframes; fi++) {
        const int f = a.frame0 + fi;
        __syncthreads();
        const float2 *Uf = a.U + wh * f;
        for (int e = tid; e < nU; e += 256) {
            const int r = e / a.TW, c = e - r * a.TW;
            const int n = X0 - a.R + c, m = Y0 - a.R + r;
            float2 v = make_float2(0.0f, 0.0f);
            if (n >= 0 && n < W && m >= 0 && m < H) v = Uf[(long)m * W + n];
            lU[r * a.pitch + c] = v;
        }
        __syncthreads();
        float acc = 0.0f;
        for (int t = 0; t < a.ntaps; t += 4) {
            const Tap c0 = lT[t], c1 = lT[t + 1], c2 = lT[t + 2], c3 = lT[t + 3];
            const float2 u0 = lU[base + c0.off];
            const float2 u1 = lU[base + c1.off];
            const float2 u2 = lU[base + c2.off];
            const float2 u3 = lU[base + c3.off];
            acc = tap_acc_f(u0, c0, acc);
            acc = tap_acc_f(u1, c1, acc);
            acc = tap_acc_f(u2, c2, acc);
            acc = tap_acc_f(u3, c3, acc);
        }
        float C = (a.wsum + acc) * invw;
        C = apply_sens_f(C, a.sens, a.sens_int);
        // unshifted frames: the intensity is the magnified pixel itself (:356)
        const float v = inimg ? a.px[wh * f + pix] : 0.0f;
        const float t = a.int_weighting ? v * C : C;
        const float e2 = a.int_weighting ? (t * v) * C : C * C;
        if (fi == 0 && a.first_init) {
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
        s4 = t;   // single 1/(n-1) TAC2 (liveSRRF_noInterpolation.cl:369-372)
    }
    if (inimg) {
        a.out[pix] = s0;
        a.out[pix + wh] = s1;
        a.out[pix + 2 * wh] = s2;
        a.out[pix + 3 * wh] = s3;
        a.prev[pix] = s4;
    }
}

size_t fht_rgc_lds_bytes(int TH, int pitch, int ntaps)
{
    return sizeof(float2) * (size_t)TH * pitch + sizeof(Tap) * (size_t)(ntaps + 4);
}

hipError_t launch_fht_rgc(const FhtRgcArgs &a, hipStream_t s)
{
    if (a.nframes <= 0) return hipSuccess;
    const size_t lds = fht_rgc_lds_bytes(a.TH, a.pitch, a.ntaps);
    hipLaunchKernelGGL(k_fht_rgc, dim3(a.tiles_x * a.tiles_y), dim3(256), lds, s, a, a.taps,
                       a.tapk);
    return hipGetLastError();
}

// Per-pixel FHT-mode RGC for frames with drift (liveSRRF_noInterpolation.cl:260-393).

__global__ void __launch_bounds__(256) k_fht_generic(const FhtGenericArgs a)
{
    const long offset = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int W = a.W, H = a.H;
    const long wh = (long)W * H;
    if (offset >= wh) return;
    const int yM = (int)(offset / W), xM = (int)(offset - (long)yM * W);
    const int jlo = -(int)a.radius, jhi = (int)(a.radius + 1.0f);
    float s0, s1, s2, s3, s4;
    if (a.count0 == 0) {
        s0 = s1 = s2 = s3 = s4 = 0.0f;
    } else {
        s0 = a.out[offset];
        s1 = a.out[offset + wh];
        s2 = a.out[offset + 2 * wh];
        s3 = a.out[offset + 3 * wh];
        s4 = a.prev[offset];
    }
    for (int fi = 0; fi < a.nframes; fi++) {
        const int c0 = a.count0 + fi, f = a.frame0 + fi;
        const float shiftX = a.drift[c0], shiftY = a.drift[c0 + a.nFrameForSRRF];
        const float xc = (float)((xM + 0.5) + (double)shiftX);
        const float yc = (float)((yM + 0.5) + (double)shiftY);
        const float2 *Uf = a.U + wh * f;
        const int bx = (int)((double)xc - 0.5), by = (int)((double)yc - 0.5);
        float CGLH = 0.0f, sumW = 0.0f;
        for (int j = jlo; j <= jhi; j++) {
            const float vy = (float)(((float)by + (float)j) + 0.5);
            if (!(vy > 0.0f && vy < (float)H)) continue;
            for (int i = jlo; i <= jhi; i++) {
                const float vx = (float)(((float)bx + (float)i) + 0.5);
                if (!(vx > 0.0f && vx < (float)W)) continue;
                const float dx = vx - xc, dy = vy - yc;
                const float d = sqrtf(dx * dx + dy * dy);
                if (!(d != 0.0f && d <= a.tSO)) continue;
                float dw = d * expf(-(d * d) / a.tSS);
                dw = dw * dw * dw * dw;
                sumW += dw;
                const int n = bx + i, m = by + j;
                const float2 u = Uf[(long)m * W + n];
                if (u.x * dx + u.y * dy < 0.0f) CGLH += (1.0f - fabsf(u.y * dx - u.x * dy) / d) * dw;
            }
        }
        CGLH /= sumW;
        CGLH = apply_sens_f(CGLH, a.sens, a.sens_int);
        float v;
        if (shiftX == 0.0f && shiftY == 0.0f)
            v = a.px[wh * f + offset];
        else
            v = interp_global(a.px + wh * f, W, H, (float)xM + shiftX, (float)yM + shiftY);
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
        s4 = t;
    }
    a.out[offset] = s0;
    a.out[offset + wh] = s1;
    a.out[offset + 2 * wh] = s2;
    a.out[offset + 3 * wh] = s3;
    a.prev[offset] = s4;
}

hipError_t launch_fht_generic(const FhtGenericArgs &a, hipStream_t s)
{
    if (a.nframes <= 0) return hipSuccess;
    const long n = (long)a.W * a.H;
    hipLaunchKernelGGL(k_fht_generic, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace esrrf
