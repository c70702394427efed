// k_frc.hip -- Fourier ring correlation of two images (the sweep's FRC resolution,
// gui/ParametersSweep_.java:403-447): crop to the centred square, Tukey taper, zero-pad to a
// power of two, 2-D FFT of both (rows then columns, Stockham in LDS), and per-ring sums of
// Re(F1 conj F2), |F1|^2 and |F2|^2 in double.
#include <algorithm>

#include "device_fft.h"
#include "frc_internal.h"

namespace esrrf {

__global__ void k_frc_prep(const float *__restrict__ a, const float *__restrict__ b, FrcPlan p)
{
    const long NN = (long)p.N * p.N;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < 2 * NN;
         e += (long)gridDim.x * blockDim.x) {
        const int img = (int)(e / NN);
        const long r = e - img * NN;
        const int y = (int)(r / p.N), x = (int)(r - (long)y * p.N);
        float v = 0.0f;
        if (y < p.S && x < p.S) {
            const float s = (img ? b : a)[(long)(p.y0 + y) * p.W + (p.x0 + x)];
            v = (s == s ? s : 0.0f) * p.win[y] * p.win[x];   // NaN reads as 0
        }
        p.buf[e] = make_float2(v, 0.0f);
    }
}

// G rows of one image per block, in place. grid (N/G, 2)
template <int G>
__global__ void __launch_bounds__(256) k_frc_rows(FrcPlan p)
{
    extern __shared__ float2 sm[];
    const int N = p.N;
    float2 *A = sm, *B = sm + G * N, *TW = sm + 2 * G * N;
    stage_twiddles(TW, p.tw, N);
    float2 *R0 = p.buf + (long)blockIdx.y * N * N + (long)blockIdx.x * G * N;
    for (int e = threadIdx.x; e < G * N; e += blockDim.x) A[e] = R0[e];
    __syncthreads();
    const float2 *R = stockham<G>(A, B, N, p.log2N, TW);
    for (int e = threadIdx.x; e < G * N; e += blockDim.x) R0[e] = R[e];
}

// G columns of one image per block, in place, tiles in XCD-grouped order. grid (N/G, batch)
template <int G>
__global__ void __launch_bounds__(256) k_frc_cols(FrcPlan p)
{
    extern __shared__ float2 sm[];
    const int N = p.N;
    float2 *A = sm, *B = sm + G * N, *TW = sm + 2 * G * N;
    stage_twiddles(TW, p.tw, N);
    const int2 t = xcd_tile(1);
    const int a0 = t.x * G;
    float2 *F = p.buf + (long)t.y * N * N;
    for (int e = threadIdx.x; e < G * N; e += blockDim.x) {
        const int g = e % G, k = e / G;
        A[g * N + k] = F[(long)k * N + a0 + g];
    }
    __syncthreads();
    const float2 *R = stockham<G>(A, B, N, p.log2N, TW);
    for (int e = threadIdx.x; e < G * N; e += blockDim.x) {
        const int g = e % G, k = e / G;
        F[(long)k * N + a0 + g] = R[g * N + k];
    }
}

// ring of frequency (kx, ky): round(sqrt(kx^2 + ky^2)), halves up, in integers
__device__ __forceinline__ int frc_ring(int k2)
{
    if (k2 == 0) return 0;
    int r = (int)sqrtf((float)k2);
    if ((r + 1) * (r + 1) - (r + 1) + 1 <= k2) r++;
    if (r > 0 && r * r - r + 1 > k2) r--;
    return r;
}

__global__ void __launch_bounds__(256) k_frc_rings(FrcPlan p)
{
    extern __shared__ double bins[];   // [3][R]
    const int N = p.N, R = N / 2;
    for (int i = threadIdx.x; i < 3 * R; i += blockDim.x) bins[i] = 0.0;
    __syncthreads();
    const long NN = (long)N * N;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < NN;
         e += (long)gridDim.x * blockDim.x) {
        const int ky = (int)(e / N), kx = (int)(e - (long)ky * N);
        const int sy = ky < N / 2 ? ky : ky - N, sx = kx < N / 2 ? kx : kx - N;
        const int r = frc_ring(sx * sx + sy * sy);
        if (r >= R) continue;
        const float2 f1 = p.buf[e], f2 = p.buf[NN + e];
        const double re1 = f1.x, im1 = f1.y, re2 = f2.x, im2 = f2.y;
        atomicAdd(&bins[r], re1 * re2 + im1 * im2);
        atomicAdd(&bins[R + r], re1 * re1 + im1 * im1);
        atomicAdd(&bins[2 * R + r], re2 * re2 + im2 * im2);
    }
    __syncthreads();
    double *out = p.part + (long)blockIdx.x * 3 * R;
    for (int i = threadIdx.x; i < 3 * R; i += blockDim.x) out[i] = bins[i];
}

__global__ void k_frc_reduce(FrcPlan p)
{
    const int n = 3 * (p.N / 2);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int b = 0; b < p.nparts; b++) acc += p.part[(long)b * n + i];
        p.rings[i] = acc;
    }
}

template <int G>
static void frc_fft(const FrcPlan &p, int batch, hipStream_t s)
{
    const size_t lds = sizeof(float2) * ((size_t)2 * G * p.N + p.N);
    hipLaunchKernelGGL((k_frc_rows<G>), dim3(p.N / G, batch), dim3(256), lds, s, p);
    hipLaunchKernelGGL((k_frc_cols<G>), dim3(p.N / G, batch), dim3(256), lds, s, p);
}

static void fft_batch(const FrcPlan &p, int batch, hipStream_t s)
{
    switch (std::max(1, std::min(std::min(16, p.N), 1024 / p.N))) {
    case 16: frc_fft<16>(p, batch, s); break;
    case 8: frc_fft<8>(p, batch, s); break;
    case 4: frc_fft<4>(p, batch, s); break;
    case 2: frc_fft<2>(p, batch, s); break;
    default: frc_fft<1>(p, batch, s); break;
    }
}

hipError_t launch_fft2d(const FrcPlan &p, int batch, hipStream_t s)
{
    if (p.N < 4 || p.N > FRC_MAX_N || batch < 1 || batch > 65535) return hipErrorInvalidValue;
    fft_batch(p, batch, s);
    return hipGetLastError();
}

hipError_t launch_frc(const FrcPlan &p, const float *a, const float *b, hipStream_t s)
{
    if (p.N < 2 || p.N > FRC_MAX_N) return hipErrorInvalidValue;
    const long n2 = 2L * p.N * p.N;
    hipLaunchKernelGGL(k_frc_prep, dim3((unsigned)std::min<long>((n2 + 255) / 256, 8192)),
                       dim3(256), 0, s, a, b, p);
    fft_batch(p, 2, s);
    hipLaunchKernelGGL(k_frc_rings, dim3(p.nparts), dim3(256),
                       sizeof(double) * 3 * (size_t)(p.N / 2), s, p);
    hipLaunchKernelGGL(k_frc_reduce, dim3((3 * (p.N / 2) + 255) / 256), dim3(256), 0, s, p);
    return hipGetLastError();
}

}  // namespace esrrf
