// device_fft.h -- batched power-of-two FFTs in LDS, shared by the FHT upsampling and the
// Fourier ring correlation.
#pragma once
#include <hip/hip_runtime.h>

namespace esrrf {

// ---------------------------------------------------------------------------------------
// Stockham FFT (forward, e^{-i}) of G signals of length N = 2^log2N held in LDS: A holds the
// input, B is scratch; returns the buffer holding the result.  Radix-4 stages (after one
// radix-2 stage when log2N is odd): with Q = N/4 and the sub-transform length L, input j
// (k = j mod L) combines A[j + m Q] w^{m k}, w = e^{-2 pi i / (4L)}, into the 4-point DFT
// stored at 4 (j - k) + k + q L (the self-sorting form).  tw (in LDS) holds
// tw[t] = e^{-2 pi i t / N} for t < N.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int G>
__device__ __forceinline__ float2 *stockham(float2 *A, float2 *B, int N, int log2N,
                                            const float2 *tw)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    int L = 1;
    if (log2N & 1) {   // radix-2 stage, L = 1 (all twiddles 1)
        const int half = N >> 1;
        for (int e = tid; e < G * half; e += nthr) {
            const int g = e >> (log2N - 1), j = e - g * half;
            const float2 u = A[g * N + j], x = A[g * N + j + half];
            B[g * N + 2 * j] = make_float2(u.x + x.x, u.y + x.y);
            B[g * N + 2 * j + 1] = make_float2(u.x - x.x, u.y - x.y);
        }
        __syncthreads();
        float2 *t = A;
        A = B;
        B = t;
        L = 2;
    }
    const int Q = N >> 2;
    for (; L < N; L <<= 2) {
        const int tstep = N / (4 * L);
        for (int e = tid; e < G * Q; e += nthr) {
            const int g = e >> (log2N - 2), j = e - g * Q;
            const int k = j & (L - 1);
            const float2 *a = A + g * N + j;
            const float2 a0 = a[0];
            const float2 a1 = cmul(a[Q], tw[k * tstep]);
            const float2 a2 = cmul(a[2 * Q], tw[2 * k * tstep]);
            const float2 a3 = cmul(a[3 * Q], tw[3 * k * tstep]);
            const float2 t0 = make_float2(a0.x + a2.x, a0.y + a2.y);
            const float2 t1 = make_float2(a0.x - a2.x, a0.y - a2.y);
            const float2 t2 = make_float2(a1.x + a3.x, a1.y + a3.y);
            const float2 t3 = make_float2(a1.x - a3.x, a1.y - a3.y);
            float2 *o = B + g * N + 4 * (j - k) + k;
            o[0] = make_float2(t0.x + t2.x, t0.y + t2.y);
            o[L] = make_float2(t1.x + t3.y, t1.y - t3.x);        // t1 - i t3
            o[2 * L] = make_float2(t0.x - t2.x, t0.y - t2.y);
            o[3 * L] = make_float2(t1.x - t3.y, t1.y + t3.x);    // t1 + i t3
        }
        __syncthreads();
        float2 *t = A;
        A = B;
        B = t;
    }
    return A;
}

// the N twiddles of the block's transforms into LDS (call before the first barrier)
__device__ __forceinline__ void stage_twiddles(float2 *TW, const float2 *__restrict__ tw, int N)
{
    for (int t = threadIdx.x; t < N; t += blockDim.x) TW[t] = tw[t];
}

// Tile of this block in a (gx, gy) grid.  Workgroups go to the eight XCDs in turn
// (workgroup i to XCD i mod 8), each XCD with its own L2.  A column pass of the Stockham FFTs
// reads G adjacent columns per block, a part of each 128-byte line, so in launch order the
// neighbouring column groups that share a line run on different XCDs and each XCD fetches
// the line.  With xcd set (and a grid that divides by eight) XCD j takes the contiguous tile
// range [j total/8, (j + 1) total/8) in order instead, and neighbours share the line in one
// L2.  Either order visits every tile once.
__device__ __forceinline__ int2 xcd_tile(int xcd)
{
    const int gx = gridDim.x, total = gx * gridDim.y;
    int lin = blockIdx.y * gx + blockIdx.x;
    if (xcd && (total & 7) == 0) lin = (lin & 7) * (total >> 3) + (lin >> 3);
    const int ty = lin / gx;
    return make_int2(lin - ty * gx, ty);
}

}  // namespace esrrf
