// k_drift.hip -- drift of frames against a reference frame (XYShiftCalculator.java:25-47):
// mean-free frames zero-padded to a power-of-two square, circular cross-correlation by FFT,
// integer peak and a three-point parabola per axis.
#include <algorithm>

#include "drift_internal.h"
#include "frc_internal.h"

namespace esrrf {

// segment blockIdx.x of frame blockIdx.y: sum of its pixels (NaN as 0) in double, into
// part[frame][segment]; grid (DRIFT_SEG, frames)
__global__ void __launch_bounds__(256) k_drift_sums(const float *__restrict__ frames, int w,
                                                     int h, double *__restrict__ part)
{
    __shared__ double red[256];
    const long wh = (long)w * h;
    const long seg = (wh + DRIFT_SEG - 1) / DRIFT_SEG;
    const long e0 = seg * blockIdx.x, e1 = e0 + seg < wh ? e0 + seg : wh;
    const float *f = frames + wh * blockIdx.y;
    double acc = 0.0;
    for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const float v = f[e];
        acc += v == v ? (double)v : 0.0;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)blockIdx.y * DRIFT_SEG + blockIdx.x] = red[0];
}

// frame sums from the partials, segments in order; one thread per frame
__global__ void k_drift_sums_fin(const double *__restrict__ part, int batch,
                                 double *__restrict__ sums)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    double acc = 0.0;
    for (int k = 0; k < DRIFT_SEG; k++) acc += part[(long)b * DRIFT_SEG + k];
    sums[b] = acc;
}

// frame b of the batch minus its mean into the top-left of an N x N complex image;
// grid (N * ceil(N / 256), frames): a block takes 256 consecutive elements of one row
__global__ void __launch_bounds__(256) k_drift_prep(const float *__restrict__ frames,
                                                     const double *__restrict__ sums, int w,
                                                     int h, int N, float2 *__restrict__ buf)
{
    const int b = blockIdx.y;
    const int cw = (N + 255) >> 8;
    const int y = blockIdx.x / cw;
    const int x = (blockIdx.x - y * cw) * 256 + threadIdx.x;
    if (x >= N) return;
    const long wh = (long)w * h;
    float v = 0.0f;
    if (x < w && y < h) {
        const float s = frames[wh * b + (long)y * w + x];
        v = (float)((double)(s == s ? s : 0.0f) - sums[b] / (double)wh);
    }
    buf[(long)N * N * b + (long)y * N + x] = make_float2(v, 0.0f);
}

// R(k) conj(F_b(k)): the conjugate of conj(R) F_b, whose inverse transform is the
// correlation; a forward transform of this product has the correlation (times N^2) as its
// real part
__global__ void k_drift_xspec(float2 *__restrict__ buf, const float2 *__restrict__ R, int N,
                              int batch)
{
    const long NN = (long)N * N;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < batch * NN;
         e += (long)gridDim.x * blockDim.x) {
        const float2 r = R[e % NN], f = buf[e];
        buf[e] = make_float2(r.x * f.x + r.y * f.y, r.y * f.x - r.x * f.y);
    }
}

// first maximum (largest value, then smallest row-major index) of the correlation over
// segment blockIdx.x of frame blockIdx.y; grid (DRIFT_SEG, frames)
__global__ void __launch_bounds__(256) k_drift_peak_seg(const float2 *__restrict__ buf, int N,
                                                         float *__restrict__ pkv,
                                                         int *__restrict__ pki)
{
    __shared__ float bv[256];
    __shared__ int bi[256];
    const int NN = N * N;
    const int seg = (NN + DRIFT_SEG - 1) / DRIFT_SEG;
    const int e0 = seg * blockIdx.x, e1 = min(e0 + seg, NN);
    const float2 *c = buf + (long)NN * blockIdx.y;
    float best = -3.4028235e38f;
    int bidx = 0x7fffffff;
    for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const float v = c[e].x;
        if (v > best) {   // ascending e per thread: keeps the thread's first maximum
            best = v;
            bidx = e;
        }
    }
    bv[threadIdx.x] = best;
    bi[threadIdx.x] = bidx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const float v = bv[threadIdx.x + s];
            const int i = bi[threadIdx.x + s];
            if (v > bv[threadIdx.x] || (v == bv[threadIdx.x] && i < bi[threadIdx.x])) {
                bv[threadIdx.x] = v;
                bi[threadIdx.x] = i;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pkv[(long)blockIdx.y * DRIFT_SEG + blockIdx.x] = bv[0];
        pki[(long)blockIdx.y * DRIFT_SEG + blockIdx.x] = bi[0];
    }
}

// per frame (one thread each): the first maximum over the segments' candidates (the same
// rule, so the same peak as one scan of the whole map), its lags wrapped to [-N/2, N/2) and
// the parabola through the peak and its neighbours per axis
__global__ void k_drift_peak(const float2 *__restrict__ buf, int N, int batch,
                             const float *__restrict__ pkv, const int *__restrict__ pki,
                             float *__restrict__ sx, float *__restrict__ sy)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    float best = pkv[(long)b * DRIFT_SEG];
    int bidx = pki[(long)b * DRIFT_SEG];
    for (int k = 1; k < DRIFT_SEG; k++) {
        const float v = pkv[(long)b * DRIFT_SEG + k];
        const int i = pki[(long)b * DRIFT_SEG + k];
        if (v > best || (v == best && i < bidx)) {
            best = v;
            bidx = i;
        }
    }
    // a map without a finite maximum (a frame holding +-Inf makes it NaN or Inf) leaves
    // bidx unset: no shift, and no read through it
    if (bidx < 0 || bidx >= N * N || !isfinite(best)) {
        sx[b] = 0.0f;
        sy[b] = 0.0f;
        return;
    }
    const float2 *c = buf + (long)N * N * b;
    const int py = bidx / N, px = bidx - py * N;
    const float c0 = best;
    const float xm = c[py * N + ((px - 1) & (N - 1))].x, xp = c[py * N + ((px + 1) & (N - 1))].x;
    const float ym = c[((py - 1) & (N - 1)) * N + px].x, yp = c[((py + 1) & (N - 1)) * N + px].x;
    const float dx_den = xm - 2.0f * c0 + xp, dy_den = ym - 2.0f * c0 + yp;
    float fx = dx_den != 0.0f ? 0.5f * (xm - xp) / dx_den : 0.0f;
    float fy = dy_den != 0.0f ? 0.5f * (ym - yp) / dy_den : 0.0f;
    if (!isfinite(fx)) fx = 0.0f;   // a non-finite neighbour: integer peak on that axis
    if (!isfinite(fy)) fy = 0.0f;
    const int lx = px < N / 2 ? px : px - N, ly = py < N / 2 ? py : py - N;
    sx[b] = (float)lx + fx;
    sy[b] = (float)ly + fy;
}

static unsigned grid_for(long n)
{
    return (unsigned)std::min<long>((n + 255) / 256, 8192);
}

hipError_t launch_drift_reference(const DriftPlan &d, const float *ref, hipStream_t s)
{
    hipLaunchKernelGGL(k_drift_sums, dim3(DRIFT_SEG, 1), dim3(256), 0, s, ref, d.w, d.h,
                       d.sums + d.max_batch);
    hipLaunchKernelGGL(k_drift_sums_fin, dim3(1), dim3(64), 0, s, d.sums + d.max_batch, 1,
                       d.sums);
    const unsigned rows = (unsigned)d.N * ((d.N + 255) / 256);
    hipLaunchKernelGGL(k_drift_prep, dim3(rows, 1), dim3(256), 0, s, ref, d.sums, d.w, d.h, d.N,
                       d.R);
    FrcPlan p{};
    p.N = d.N;
    p.log2N = d.log2N;
    p.tw = d.tw;
    p.buf = d.R;
    return launch_fft2d(p, 1, s);
}

hipError_t launch_drift_batch(const DriftPlan &d, const float *frames, int batch, float *sx,
                              float *sy, hipStream_t s)
{
    if (batch < 1 || batch > d.max_batch) return hipErrorInvalidValue;
    const long n = (long)batch * d.N * d.N;
    hipLaunchKernelGGL(k_drift_sums, dim3(DRIFT_SEG, batch), dim3(256), 0, s, frames, d.w, d.h,
                       d.sums + d.max_batch);
    hipLaunchKernelGGL(k_drift_sums_fin, dim3((batch + 63) / 64), dim3(64), 0, s,
                       d.sums + d.max_batch, batch, d.sums);
    const unsigned rows = (unsigned)d.N * ((d.N + 255) / 256);
    hipLaunchKernelGGL(k_drift_prep, dim3(rows, batch), dim3(256), 0, s, frames, d.sums, d.w,
                       d.h, d.N, d.buf);
    FrcPlan p{};
    p.N = d.N;
    p.log2N = d.log2N;
    p.tw = d.tw;
    p.buf = d.buf;
    hipError_t e = launch_fft2d(p, batch, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_drift_xspec, dim3(grid_for(n)), dim3(256), 0, s, d.buf, d.R, d.N,
                       batch);
    e = launch_fft2d(p, batch, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_drift_peak_seg, dim3(DRIFT_SEG, batch), dim3(256), 0, s, d.buf, d.N,
                       d.pkv, d.pki);
    hipLaunchKernelGGL(k_drift_peak, dim3((batch + 63) / 64), dim3(64), 0, s, d.buf, d.N, batch,
                       d.pkv, d.pki, sx, sy);
    return hipGetLastError();
}

}  // namespace esrrf
