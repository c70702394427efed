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
// with k() the signed frequency.  The twiddle is separable, t_b(kx) t_d(ky), so the row
// transforms are shared by the M values of d: per frame one forward N x N FFT, M row passes
// and M^2 column passes of N-point FFTs, all power of two (radix-2 Stockham in LDS, G signals
// per block), batched over chunks of frames, with the mirror padding fused into the first
// load and the (Re - Im) crop-scatter into the last store.
#include "device_fft.h"
#include "device_interp.h"
#include "device_tiles.h"
#include "esrrf_internal.h"
#include "fht_internal.h"

#include <algorithm>
#include <cstdlib>

namespace esrrf {

// twiddle e^{-2 pi i k(q) v / (MN)} of frequency index q (signed frequency k(q), Nyquist on
// the negative side) from the table twMN[j] = e^{-2 pi i j / (MN)}
__device__ __forceinline__ float2 twid(const float2 *__restrict__ twMN, int q, int v, int N, int MN)
{
    const int k = q < N / 2 ? q : q - N;
    int j = (k * v) % MN;
    if (j < 0) j += MN;
    return twMN[j];
}

// Step 1: forward FFT along x of G rows of the mirror-padded frame
// (FourierInterpolation.java:118-172; the identity for power-of-two squares).
// grid (N/G, frames); F[f][y][kx]
template <int G>
__global__ void __launch_bounds__(256) k_fht_fwd_rows(const float *__restrict__ src, long src_stride,
                                                      float2 *__restrict__ F, FhtPlan p)
{
    extern __shared__ float2 sm[];
    const int N = p.N;
    float2 *A = sm, *B = sm + G * N, *TW = sm + 2 * G * N;
    stage_twiddles(TW, p.tw, N);
    const int y0 = blockIdx.x * G, f = blockIdx.y;
    const float *S = src + src_stride * f;
    for (int e = threadIdx.x; e < G * N; e += blockDim.x) {
        const int g = e / N, xx = e - g * N;
        const int X = xx + p.xROIp, Y = y0 + g + p.yROIp;
        const int sW = X / p.w, i = X - sW * p.w;
        const int sH = Y / p.h, j = Y - sH * p.h;
        const int px = ((sW - (p.scaleW - 1) / 2) % 2 == 0) ? i : p.w - 1 - i;
        const int qy = ((sH - (p.scaleH - 1) / 2) % 2 == 0) ? j : p.h - 1 - j;
        A[e] = make_float2(S[(long)qy * p.w + px], 0.0f);
    }
    __syncthreads();
    const float2 *R = stockham<G>(A, B, N, p.log2N, TW);
    float2 *Fo = F + (long)N * N * f + (long)y0 * N;
    for (int e = threadIdx.x; e < G * N; e += blockDim.x) Fo[e] = R[e];
}

// Step 2: forward FFT along y of G columns, in place. grid (N/G, frames)
template <int G>
__global__ void __launch_bounds__(256) k_fht_fwd_cols(float2 *__restrict__ F, FhtPlan p)
{
    extern __shared__ float2 sm[];
    const int N = p.N;
    float2 *A = sm, *B = sm + G * N, *TW = sm + 2 * G * N;
    stage_twiddles(TW, p.tw, N);
    const int2 t = xcd_tile(p.xcd);
    const int a0 = t.x * G;
    float2 *Ff = F + (long)N * N * t.y;
    for (int e = threadIdx.x; e < G * N; e += blockDim.x) {
        const int g = e % G, k = e / G;          // consecutive threads: consecutive columns
        A[g * N + k] = Ff[(long)k * N + a0 + g];
    }
    __syncthreads();
    const float2 *R = stockham<G>(A, B, N, p.log2N, TW);
    for (int e = threadIdx.x; e < G * N; e += blockDim.x) {
        const int g = e % G, k = e / G;
        Ff[(long)k * N + a0 + g] = R[g * N + k];
    }
}

// Step 3: Hartley spectrum H = Re F - Im F of G rows (ky), twiddled by t_b(kx), FFT along
// kx -> R[f][b][ky][a].  grid (N/G, frames * M), b = blockIdx.y % M
template <int G>
__global__ void __launch_bounds__(256) k_fht_inv_rows(const float2 *__restrict__ F,
                                                      float2 *__restrict__ Rb, FhtPlan p)
{
    extern __shared__ float2 sm[];
    const int N = p.N, M = p.M, MN = p.M * p.N;
    float2 *A = sm, *B = sm + G * N, *TW = sm + 2 * G * N;
    stage_twiddles(TW, p.tw, N);
    const int y0 = blockIdx.x * G;
    const int f = blockIdx.y / M, b = blockIdx.y - f * M;
    const float2 *Fr = F + (long)N * N * f + (long)y0 * N;
    for (int e = threadIdx.x; e < G * N; e += blockDim.x) {
        const int kx = e % N;
        const float2 v = Fr[e];
        const float hv = v.x - v.y;
        const float2 t = twid(p.twMN, kx, b, N, MN);
        A[e] = make_float2(hv * t.x, hv * t.y);
    }
    __syncthreads();
    const float2 *R = stockham<G>(A, B, N, p.log2N, TW);
    float2 *Ro = Rb + ((long)f * M + b) * N * N + (long)y0 * N;
    for (int e = threadIdx.x; e < G * N; e += blockDim.x) Ro[e] = R[e];
}

// Step 4: column a of R_b twiddled by t_d(ky), FFT along ky: element c is the inverse-DHT
// sample at (M c + d, M a + b) of the MN x MN grid; (Re - Im)/(MN)^2 cropped into the
// magnified frame.  grid (N/G, frames * M * M), (b, d) from blockIdx.y
template <int G>
__global__ void __launch_bounds__(256) k_fht_inv_cols(const float2 *__restrict__ Rb,
                                                      float *__restrict__ out, long out_stride,
                                                      FhtPlan p, float inv)
{
    extern __shared__ float2 sm[];
    const int N = p.N, M = p.M, MN = p.M * p.N;
    const int wM = p.M * p.w, hM = p.M * p.h;
    float2 *A = sm, *B = sm + G * N, *TW = sm + 2 * G * N;
    stage_twiddles(TW, p.tw, N);
    const int2 tl = xcd_tile(p.xcd);
    const int a0 = tl.x * G;
    const int MM = M * M;
    const int f = tl.y / MM, bd = tl.y - f * MM;
    const int b = bd / M, d = bd - b * M;
    const float2 *Rc = Rb + ((long)f * M + b) * N * N;
    for (int e = threadIdx.x; e < G * N; e += blockDim.x) {
        const int g = e % G, k = e / G;
        const float2 v = Rc[(long)k * N + a0 + g];
        const float2 t = twid(p.twMN, k, d, N, MN);
        A[g * N + k] = make_float2(v.x * t.x - v.y * t.y, v.x * t.y + v.y * t.x);
    }
    __syncthreads();
    const float2 *R = stockham<G>(A, B, N, p.log2N, TW);
    float *O = out + out_stride * f;
    for (int e = threadIdx.x; e < G * N; e += blockDim.x) {
        const int g = e % G, c = e / G;
        const int X = M * (a0 + g) + b - p.xROI, Y = M * c + d - p.yROI;
        if (X < 0 || X >= wM || Y < 0 || Y >= hM) continue;
        const float2 z = R[g * N + c];
        O[(long)Y * wM + X] = (z.x - z.y) * inv;
    }
}

// Step 4 with coalesced stores: the block takes column group a0..a0+G-1 of one (frame, d)
// and runs the column transforms of every b in turn (each exactly as k_fht_inv_cols does),
// collecting (Re - Im)/(MN)^2 in an LDS tile of output rows y = M c + d, x = M a + b; the
// tile then goes out as runs of M G consecutive pixels per row instead of pixels M apart.
// With PF (G N <= 1024 and 256 threads) each thread holds its elements of the next b's
// columns in registers, loaded while this b's transform runs in LDS.
// grid (N/G, frames * M), d from blockIdx.y
template <int G, bool PF>
__global__ void __launch_bounds__(256) k_fht_inv_cols_rows(const float2 *__restrict__ Rb,
                                                           float *__restrict__ out,
                                                           long out_stride, FhtPlan p, float inv)
{
    extern __shared__ float2 sm[];
    const int N = p.N, M = p.M, MN = p.M * p.N;
    const int wM = p.M * p.w, hM = p.M * p.h;
    float2 *A = sm, *B = sm + G * N, *TW = sm + 2 * G * N;
    float *T = reinterpret_cast<float *>(TW + N);   // [N][M G]
    stage_twiddles(TW, p.tw, N);
    const int MG = M * G;
    const int2 tl = xcd_tile(p.xcd);
    const int a0 = tl.x * G;
    const int f = tl.y / M, d = tl.y - f * M;
    float2 nv[4];
    auto fetch = [&](int b) {
        const float2 *Rc = Rb + ((long)f * M + b) * N * N;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int e = threadIdx.x + 256 * i, g = e % G, k = e / G;
            if (e < G * N) nv[i] = Rc[(long)k * N + a0 + g];
        }
    };
    if (PF) fetch(0);
    for (int b = 0; b < M; b++) {
        if (PF) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int e = threadIdx.x + 256 * i, g = e % G, k = e / G;
                if (e >= G * N) break;
                const float2 v = nv[i];
                const float2 t = twid(p.twMN, k, d, N, MN);
                A[g * N + k] = make_float2(v.x * t.x - v.y * t.y, v.x * t.y + v.y * t.x);
            }
        } else {
            const float2 *Rc = Rb + ((long)f * M + b) * N * N;
            for (int e = threadIdx.x; e < G * N; e += blockDim.x) {
                const int g = e % G, k = e / G;
                const float2 v = Rc[(long)k * N + a0 + g];
                const float2 t = twid(p.twMN, k, d, N, MN);
                A[g * N + k] = make_float2(v.x * t.x - v.y * t.y, v.x * t.y + v.y * t.x);
            }
        }
        __syncthreads();
        if (PF && b + 1 < M) fetch(b + 1);
        const float2 *R = stockham<G>(A, B, N, p.log2N, TW);
        for (int e = threadIdx.x; e < G * N; e += blockDim.x) {
            const int g = e % G, c = e / G;
            const float2 z = R[g * N + c];
            T[c * MG + M * g + b] = (z.x - z.y) * inv;
        }
        __syncthreads();   // A and B are reused by the next b
    }
    float *O = out + out_stride * f;
    const int x0 = M * a0 - p.xROI;
    for (int e = threadIdx.x; e < N * MG; e += blockDim.x) {
        const int c = e / MG, xx = e - c * MG;
        const int X = x0 + xx, Y = M * c + d - p.yROI;
        if (X < 0 || X >= wM || Y < 0 || Y >= hM) continue;
        O[(long)Y * wM + X] = T[e];
    }
}

static int grid_of(long n)
{
    long g = (n + 255) / 256;
    return (int)(g > 4096 ? 4096 : (g < 1 ? 1 : g));
}

// The row-tiled last pass: G columns per block with A, B, the twiddles and the output tile
// within 40 KiB (measured at N = 256, M = 5: G = 4 runs the upsampling of 341 frames in
// 11.8 ms, G = 8 in 15.7 ms, G = 1 in 14.9 ms).
template <int G>
static void inv_cols_rows_g(const FhtPlan &p, const float2 *Y, float *out, long stride,
                            float inv, int k, hipStream_t s)
{
    const size_t lds = sizeof(float2) * ((size_t)2 * G * p.N + p.N) +
                       sizeof(float) * (size_t)p.N * p.M * G;
    const dim3 grid(p.N / G, k * p.M);
    if (G * p.N <= 1024 && !p.no_prefetch)
        hipLaunchKernelGGL((k_fht_inv_cols_rows<G, true>), grid, dim3(256), lds, s, Y, out,
                           stride, p, inv);
    else
        hipLaunchKernelGGL((k_fht_inv_cols_rows<G, false>), grid, dim3(256), lds, s, Y, out,
                           stride, p, inv);
}

static void inv_cols_rows(const FhtPlan &p, const float2 *Y, float *out, long stride, float inv,
                          int k, hipStream_t s)
{
    const size_t per = 2 * sizeof(float2) * (size_t)p.N + sizeof(float) * (size_t)p.N * p.M;
    int G = 16;
    while (G > 1 && (G > p.N || (size_t)G * per + sizeof(float2) * p.N > 40 * 1024)) G >>= 1;
    switch (G) {
    case 16: inv_cols_rows_g<16>(p, Y, out, stride, inv, k, s); break;
    case 8: inv_cols_rows_g<8>(p, Y, out, stride, inv, k, s); break;
    case 4: inv_cols_rows_g<4>(p, Y, out, stride, inv, k, s); break;
    case 2: inv_cols_rows_g<2>(p, Y, out, stride, inv, k, s); break;
    default: inv_cols_rows_g<1>(p, Y, out, stride, inv, k, s); break;
    }
}

template <int G>
static void upsample_g(const FhtPlan &p, const float *frames, float *out, int nf, hipStream_t s)
{
    const int N = p.N, MM = p.M * p.M;
    const size_t lds = sizeof(float2) * ((size_t)2 * G * N + N);
    const long npx = (long)p.w * p.h, nM = (long)MM * npx;
    const float inv = (float)(1.0 / ((double)p.M * p.N * (double)p.M * p.N));
    for (int f0 = 0; f0 < nf; f0 += p.chunk) {
        const int k = std::min(p.chunk, nf - f0);
        hipLaunchKernelGGL((k_fht_fwd_rows<G>), dim3(N / G, k), dim3(256), lds, s,
                           frames + npx * f0, npx, p.F, p);
        hipLaunchKernelGGL((k_fht_fwd_cols<G>), dim3(N / G, k), dim3(256), lds, s, p.F, p);
        hipLaunchKernelGGL((k_fht_inv_rows<G>), dim3(N / G, k * p.M), dim3(256), lds, s, p.F,
                           p.Y, p);
        if (p.scatter)
            hipLaunchKernelGGL((k_fht_inv_cols<G>), dim3(N / G, k * MM), dim3(256), lds, s,
                               p.Y, out + nM * f0, nM, p, inv);
        else
            inv_cols_rows(p, p.Y, out + nM * f0, nM, inv, k, s);
    }
}

int fht_group(int N)
{
    // G signals per block (G <= N so the grid N / G is whole): about 16 KiB of signals per
    // block, so that several blocks share a CU (measured at N = 256: G = 4 and 8 run the
    // passes in 15.6 ms per 341 frames, G = 16 in 16.7 ms)
    return std::max(1, std::min(std::min(16, N), 1024 / N));
}

// Upsample nf consecutive raw frames (float, w x h each) into nf magnified frames.
hipError_t launch_fht_upsample(const FhtPlan &p, const float *frames, float *out, int nf,
                               hipStream_t s)
{
    if (nf <= 0) return hipSuccess;
    switch (fht_group(p.N)) {
    case 16: upsample_g<16>(p, frames, out, nf, s); break;
    case 8: upsample_g<8>(p, frames, out, nf, s); break;
    case 4: upsample_g<4>(p, frames, out, nf, s); break;
    case 2: upsample_g<2>(p, frames, out, nf, s); break;
    default: upsample_g<1>(p, frames, out, nf, s); break;
    }
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

// Block: 64 x 16 magnified pixels of one frame (frames over grid.z).  A pixel's unit vector
// folds the two clamped gradient reads of liveSRRF_noInterpolation.cl:291-292,
// Gx[m][min(n+1, W-1)] and Gy[min(m+1, H-1)][n], each a Roberts cross of the 2 x 2 pixels
// ending at its point with edge-clamped neighbours; together they reach columns n-1 .. n+1
// and rows m-1 .. m+1.  The block stages those pixels (edge-clamped, the tile origin at
// (X0 - 1, Y0 - 1)) in LDS and evaluates robx on the tile with robx's own coordinates.
constexpr int FP_X = 64, FP_Y = 16, FP_TW = FP_X + 2, FP_TH = FP_Y + 2;

__global__ void __launch_bounds__(256) k_fht_prep(const float *__restrict__ px,
                                                  float2 *__restrict__ U, int W, int H, int nF)
{
    __shared__ float tile[FP_TH * FP_TW];
    const int X0 = blockIdx.x * FP_X, Y0 = blockIdx.y * FP_Y;
    const int tid = threadIdx.x, tx = tid & (FP_X - 1), ty = tid >> 6;
    const long wh = (long)W * H;
    // pixel (x, y) of the image at tile[(y - Y0 + 1)][(x - X0 + 1)]
    auto at = [&](int x, int y) { return tile[(y - Y0 + 1) * FP_TW + (x - X0 + 1)]; };
#pragma unroll 1
    for (int f = blockIdx.z; f < nF; f += gridDim.z) {
        const float *P = px + wh * f;
        __syncthreads();   // the previous frame's tile is no longer read
        for (int e = tid; e < FP_TH * FP_TW; e += 256) {
            const int r = e / FP_TW, c = e - r * FP_TW;
            const int x = min(max(X0 - 1 + c, 0), W - 1), y = min(max(Y0 - 1 + r, 0), H - 1);
            tile[e] = P[(long)y * W + x];
        }
        __syncthreads();
        const int n = X0 + tx;
        if (n >= W) continue;
#pragma unroll
        for (int k = 0; k < FP_Y / 4; k++) {
            const int m = Y0 + ty + 4 * k;
            if (m >= H) break;
            float gx, gy;
            {   // robx at (min(n+1, W-1), m), its x component
                const int x1 = min(n + 1, W - 1), y1 = m;
                const int x0 = x1 - 1 < 0 ? 0 : x1 - 1, y0 = y1 - 1 < 0 ? 0 : y1 - 1;
                const float p01 = at(x1, y0), p10 = at(x0, y1), p11 = at(x1, y1), p00 = at(x0, y0);
                gx = ((p01 - p10) + p11) - p00;
            }
            {   // robx at (n, min(m+1, H-1)), its y component
                const int x1 = n, y1 = min(m + 1, H - 1);
                const int x0 = x1 - 1 < 0 ? 0 : x1 - 1, y0 = y1 - 1 < 0 ? 0 : y1 - 1;
                const float p01 = at(x1, y0), p10 = at(x0, y1), p11 = at(x1, y1), p00 = at(x0, y0);
                gy = ((-p01 + p10) + p11) - p00;
            }
            U[wh * f + (long)m * W + n] = unit_or_zero_f(gx, gy);
        }
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
    if (nF <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fht_prep, dim3((W + FP_X - 1) / FP_X, (H + FP_Y - 1) / FP_Y,
                                        std::min(nF, 65535)),
                       dim3(256), 0, s, px, U, W, H, nF);
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

// Temporal accumulation of one pixel and frame (liveSRRF_noInterpolation.cl:358-372) with
// every rounding spelled out, so that the kernels below agree bit for bit whatever the
// compiler contracts.
__device__ __forceinline__ void cumulate_f(float &s0, float &s1, float &s2, float &s3, float &s4,
                                           float C0, float v, bool first, const FhtRgcArgs &a)
{
    const float C = apply_sens_f(C0, a.sens, a.sens_int);
    const float t = a.int_weighting ? __fmul_rn(v, C) : C;
    const float e2 = a.int_weighting ? __fmul_rn(__fmul_rn(t, v), C) : __fmul_rn(C, C);
    if (first) {
        s0 = __fmul_rn(t, a.rn);
        s1 = __fmul_rn(e2, a.rn);
        s2 = 0.0f;
        s3 = __fmul_rn(v, a.rn);
    } else {
        s0 = fmaf(t, a.rn, s0);
        s1 = fmaf(e2, a.rn, s1);
        s2 = fmaf(__fmul_rn(t, s4), a.rn1, s2);
        s3 = fmaf(v, a.rn, s3);
    }
    s4 = t;   // single 1/(n-1) TAC2 (liveSRRF_noInterpolation.cl:369-372)
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
    const int lin = xcd_tile_1d(a.tiles_x * a.tiles_y);
    const int tX = lin % a.tiles_x, tY = lin / a.tiles_x;
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
    if (!a.first_init && inimg && !a.emitC) {
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
        if (a.emitC) {
            if (inimg) a.emitC[wh * f + pix] = (a.wsum + acc) * invw;
            continue;
        }
        // unshifted frames: the intensity is the magnified pixel itself (:356)
        const float v = inimg ? a.px[wh * f + pix] : 0.0f;
        cumulate_f(s0, s1, s2, s3, s4, (a.wsum + acc) * invw, v, fi == 0 && a.first_init, a);
    }
    if (inimg && !a.emitC) {
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

// ---------------------------------------------------------------------------------------
// Batched variant: each lane takes two pixels (x and x+8) of two consecutive frames.  The
// block (8 waves) covers 32 x 32 pixels; the U tile of the frame pair is interleaved as
// float4 (Ux, Uy of f, Ux, Uy of f+1) with a pitch of 8 (mod 16) float4, so one uniform
// tap read and one address add serve four samples and both 16-byte reads of a tap sit at
// compile-time offsets.  Per pixel the frames are accumulated in order with the same
// operations as k_fht_rgc: the results are equal bit for bit.
// ---------------------------------------------------------------------------------------
constexpr int FT2 = 32;

__host__ __device__ __forceinline__ int fht2_pitch(int TW2)
{
    int p = 8;
    while (p < TW2) p += 16;
    return p;
}

size_t fht_rgc2_lds_bytes(const FhtRgcArgs &a)
{
    const int TW2 = FT2 + a.R + a.Rhi;
    return sizeof(float4) * (size_t)TW2 * fht2_pitch(TW2) + sizeof(Tap) * (size_t)(a.ntaps + 4);
}

__global__ void __launch_bounds__(512) k_fht_rgc2(const FhtRgcArgs a,
                                                  const Tap *__restrict__ taps,
                                                  const int *__restrict__ tapk)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int TW2 = FT2 + a.R + a.Rhi;
    const int P2 = fht2_pitch(TW2);
    float4 *lU = reinterpret_cast<float4 *>(smem);
    Tap *lT = reinterpret_cast<Tap *>(lU + TW2 * P2);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lx = (lane & 7) + 16 * (wave & 1), ly = (lane >> 3) + 8 * (wave >> 1);
    const int lin = xcd_tile_1d(a.tiles_x * a.tiles_y);
    const int tX = lin % a.tiles_x, tY = lin / a.tiles_x;
    const int X0 = tX * FT2, Y0 = tY * FT2;
    const int xa = X0 + lx, xb = xa + 8, yM = Y0 + ly;
    const int W = a.W, H = a.H;
    const bool ina = xa < W && yM < H, inb = xb < W && yM < H;
    const long wh = (long)W * H;
    // tap table with offsets for this tile's pitch (dx, dy from tapk)
    for (int e = tid; e < a.ntaps + 4; e += 512) {
        Tap t = taps[e];
        const int k = tapk[e];
        t.off = (((k >> 8) & 255) - 128) * P2 + ((k & 255) - 128);
        lT[e] = t;
    }
    const bool interior = X0 - a.R >= 0 && X0 + FT2 - 1 + a.Rhi <= W - 1 && Y0 - a.R >= 0 &&
                          Y0 + FT2 - 1 + a.Rhi <= H - 1;
    float wsa = a.wsum, wsb = a.wsum;
    if (!interior) {
        wsa = 0.0f;
        wsb = 0.0f;
        for (int t = 0; t < a.ntaps; t++) {
            const int k = tapk[t];
            const int dn = (k & 255) - 128, m = yM + ((k >> 8) & 255) - 128;
            const bool my = m >= 0 && m <= H - 1;
            if (my && xa + dn >= 0 && xa + dn <= W - 1) wsa += taps[t].W;
            if (my && xb + dn >= 0 && xb + dn <= W - 1) wsb += taps[t].W;
        }
    }
    const float invwa = 1.0f / wsa, invwb = 1.0f / wsb;
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f, a4 = 0.0f;
    float b0 = 0.0f, b1 = 0.0f, b2 = 0.0f, b3 = 0.0f, b4 = 0.0f;
    const long pixa = (long)yM * W + xa, pixb = pixa + 8;
    if (!a.first_init && !a.emitC) {
        if (ina) {
            a0 = a.out[pixa];
            a1 = a.out[pixa + wh];
            a2 = a.out[pixa + 2 * wh];
            a3 = a.out[pixa + 3 * wh];
            a4 = a.prev[pixa];
        }
        if (inb) {
            b0 = a.out[pixb];
            b1 = a.out[pixb + wh];
            b2 = a.out[pixb + 2 * wh];
            b3 = a.out[pixb + 3 * wh];
            b4 = a.prev[pixb];
        }
    }
    const int base = (ly + a.R) * P2 + (lx + a.R);
    const int nU = TW2 * TW2;
    for (int fi = 0; fi < a.nframes; fi += 2) {
        const int f = a.frame0 + fi;
        const bool two = fi + 1 < a.nframes;
        __syncthreads();
        const float2 *Uf = a.U + wh * f;
        for (int e = tid; e < nU; e += 512) {
            const int r = e / TW2, c = e - r * TW2;
            const int n = X0 - a.R + c, m = Y0 - a.R + r;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (n >= 0 && n < W && m >= 0 && m < H) {
                const long g = (long)m * W + n;
                const float2 u0 = Uf[g];
                const float2 u1 = two ? Uf[g + wh] : make_float2(0.0f, 0.0f);
                v = make_float4(u0.x, u0.y, u1.x, u1.y);
            }
            lU[r * P2 + c] = v;
        }
        __syncthreads();
        float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, acc3 = 0.0f;
        for (int t = 0; t < a.ntaps; t += 4) {
            const Tap c0 = lT[t], c1 = lT[t + 1], c2 = lT[t + 2], c3 = lT[t + 3];
            const float4 *q0 = lU + base + c0.off;
            const float4 *q1 = lU + base + c1.off;
            const float4 *q2 = lU + base + c2.off;
            const float4 *q3 = lU + base + c3.off;
            const float4 u0 = q0[0], v0 = q0[8];
            const float4 u1 = q1[0], v1 = q1[8];
            const float4 u2 = q2[0], v2 = q2[8];
            const float4 u3 = q3[0], v3 = q3[8];
            acc0 = tap_acc_f(make_float2(u0.x, u0.y), c0, acc0);
            acc1 = tap_acc_f(make_float2(u0.z, u0.w), c0, acc1);
            acc2 = tap_acc_f(make_float2(v0.x, v0.y), c0, acc2);
            acc3 = tap_acc_f(make_float2(v0.z, v0.w), c0, acc3);
            acc0 = tap_acc_f(make_float2(u1.x, u1.y), c1, acc0);
            acc1 = tap_acc_f(make_float2(u1.z, u1.w), c1, acc1);
            acc2 = tap_acc_f(make_float2(v1.x, v1.y), c1, acc2);
            acc3 = tap_acc_f(make_float2(v1.z, v1.w), c1, acc3);
            acc0 = tap_acc_f(make_float2(u2.x, u2.y), c2, acc0);
            acc1 = tap_acc_f(make_float2(u2.z, u2.w), c2, acc1);
            acc2 = tap_acc_f(make_float2(v2.x, v2.y), c2, acc2);
            acc3 = tap_acc_f(make_float2(v2.z, v2.w), c2, acc3);
            acc0 = tap_acc_f(make_float2(u3.x, u3.y), c3, acc0);
            acc1 = tap_acc_f(make_float2(u3.z, u3.w), c3, acc1);
            acc2 = tap_acc_f(make_float2(v3.x, v3.y), c3, acc2);
            acc3 = tap_acc_f(make_float2(v3.z, v3.w), c3, acc3);
        }
        if (a.emitC) {
            float *Cf = a.emitC + wh * f;
            if (ina) Cf[pixa] = (a.wsum + acc0) * invwa;
            if (inb) Cf[pixb] = (a.wsum + acc2) * invwb;
            if (two && ina) Cf[pixa + wh] = (a.wsum + acc1) * invwa;
            if (two && inb) Cf[pixb + wh] = (a.wsum + acc3) * invwb;
            continue;
        }
        // unshifted frames: the intensity is the magnified pixel itself (:356)
        const float *Pf = a.px + wh * f;
        const float va0 = ina ? Pf[pixa] : 0.0f, vb0 = inb ? Pf[pixb] : 0.0f;
        const float va1 = (ina && two) ? Pf[pixa + wh] : 0.0f;
        const float vb1 = (inb && two) ? Pf[pixb + wh] : 0.0f;
        const bool first = fi == 0 && a.first_init;
        cumulate_f(a0, a1, a2, a3, a4, (a.wsum + acc0) * invwa, va0, first, a);
        cumulate_f(b0, b1, b2, b3, b4, (a.wsum + acc2) * invwb, vb0, first, a);
        if (two) {
            cumulate_f(a0, a1, a2, a3, a4, (a.wsum + acc1) * invwa, va1, false, a);
            cumulate_f(b0, b1, b2, b3, b4, (a.wsum + acc3) * invwb, vb1, false, a);
        }
    }
    if (a.emitC) return;
    if (ina) {
        a.out[pixa] = a0;
        a.out[pixa + wh] = a1;
        a.out[pixa + 2 * wh] = a2;
        a.out[pixa + 3 * wh] = a3;
        a.prev[pixa] = a4;
    }
    if (inb) {
        a.out[pixb] = b0;
        a.out[pixb + wh] = b1;
        a.out[pixb + 2 * wh] = b2;
        a.out[pixb + 3 * wh] = b3;
        a.prev[pixb] = b4;
    }
}

hipError_t launch_fht_rgc(const FhtRgcArgs &a, hipStream_t s, const char **kname)
{
    if (a.nframes <= 0) return hipSuccess;
    const bool pair = a.pair && fht_rgc2_lds_bytes(a) <= 65536;
    if (kname) *kname = pair ? "k_fht_rgc2" : "k_fht_rgc";
    if (pair) {
        const int tx = (a.W + FT2 - 1) / FT2, ty = (a.H + FT2 - 1) / FT2;
        FhtRgcArgs b = a;
        b.tiles_x = tx;
        b.tiles_y = ty;
        hipLaunchKernelGGL(k_fht_rgc2, dim3(tx * ty), dim3(512), fht_rgc2_lds_bytes(a), s, b,
                           a.taps, a.tapk);
        return hipGetLastError();
    }
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
