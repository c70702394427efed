// k_common.hip -- frame conversion and the read-out epilogues of the eSRRF engine.
//
//   VAR / TAC2 finalise          liveSRRF.cl:427-435, live3DSRRF.cl:524-533
//   macro-pixel (MP) map         liveSRRF.cl:464-487, live3DSRRF.cl:561-588
//   MP correction                liveSRRF.cl:490-505, live3DSRRF.cl:591-611
//   NaN -> 0 on read-back        LiveSRRF_CL.java:789
//
// All of these are HBM-bound streaming kernels over the 4 output channels.
#include <algorithm>

#include "esrrf_internal.h"

namespace esrrf {

__global__ void k_convert_u16(const uint16_t *__restrict__ src, float *__restrict__ dst, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = (float)src[i];
}

// four consecutive pixels per thread: one 8-byte load and one 16-byte store
__global__ void k_convert_u16x4(const ushort4 *__restrict__ src, float4 *__restrict__ dst, long n4)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n4; i += stride) {
        const ushort4 v = src[i];
        dst[i] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
    }
}

__global__ void k_copy_f32(const float *__restrict__ src, float *__restrict__ dst, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

static int grid_for(long n, int block)
{
    long g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (int)g;
}

hipError_t launch_convert_frames(const void *src, int dtype, float *dst, long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (dtype == 0) {
        // aligned runs of four elements as vectors, the rest (or a misaligned chunk) one by one
        long n4 = ((uintptr_t)src % 8 == 0 && (uintptr_t)dst % 16 == 0) ? n / 4 : 0;
        if (n4 > 0)
            hipLaunchKernelGGL(k_convert_u16x4, dim3(grid_for(n4, 256)), dim3(256), 0, s,
                               (const ushort4 *)src, (float4 *)dst, n4);
        const long done = 4 * n4;
        if (done < n)
            hipLaunchKernelGGL(k_convert_u16, dim3(grid_for(n - done, 256)), dim3(256), 0, s,
                               (const uint16_t *)src + done, dst + done, n - done);
    } else
        hipLaunchKernelGGL(k_copy_f32, dim3(grid_for(n, 256)), dim3(256), 0, s,
                           (const float *)src, dst, n);
    return hipGetLastError();
}

// VAR = E[x^2] - AVG^2 and TAC2 -= AVG^2, in place on channels 1 and 2.
// blockIdx.y: image of a batch at a stride of `ostride` floats
__global__ void k_var(float *__restrict__ out, long nvox, long ostride)
{
    out += ostride * blockIdx.y;
    long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; o < nvox; o += stride) {
        const float a = out[o];
        const float av2 = a * a;
        out[o + nvox] = out[o + nvox] - av2;
        out[o + 2 * nvox] = out[o + 2 * nvox] - av2;
    }
}

hipError_t launch_var(float *out, long nvox, hipStream_t s, int nbatch, long ostride)
{
    hipLaunchKernelGGL(k_var, dim3(grid_for(nvox, 256), nbatch), dim3(256), 0, s, out, nvox,
                       ostride);
    return hipGetLastError();
}

// One block per MP-map entry (channel, z, y, x sub-position): mean of that sub-position over
// every macro-pixel (M x M) or macro-voxel (M x M x Mz).  An image is nP planes of
// (w*M) x (h*M) pixels magnified to nP*Mz planes; a 2-D image is nP = Mz = 1.  The reference
// sums sequentially in float; this tree-reduces in double, which is at least as accurate.
// The block is 64 x 16 threads: thread (tx, ty) sums macro-pixel columns tx, tx + 64, ... of
// macro-pixel rows ty, ty + 16, ... over every plane, so the index arithmetic has one division
// per row and none per element.
// blockIdx.y: image of a batch (images `ostride` floats apart, maps `mstride` floats apart).
constexpr int MP_TX = 64, MP_TY = 16, MP_THREADS = MP_TX * MP_TY;

__global__ void __launch_bounds__(MP_THREADS) k_mp_map(const float *__restrict__ out,
                                                       float *__restrict__ mp, int w, int h,
                                                       int nP, int M, int Mz, long ostride,
                                                       long mstride)
{
    out += ostride * blockIdx.y;
    mp += mstride * blockIdx.y;
    const int o = blockIdx.x;
    const int MM = M * M;
    const int Md = MM * Mz;
    const int frame = o / Md;
    const int r = o - frame * Md;
    const int zMP = r / MM;
    const int yMP = (r - zMP * MM) / M;
    const int xMP = r - zMP * MM - yMP * M;
    const long wM = (long)w * M;
    const long whM = wM * h * M;
    const long whdM = whM * nP * Mz;
    const float *base = out + xMP + wM * yMP + whM * zMP + frame * whdM;
    const int tx = threadIdx.x, ty = threadIdx.y;
    double acc = 0.0;
    for (int row = ty; row < h * nP; row += MP_TY) {
        const int z = row / h, y = row - z * h;
        const float *rp = base + whM * ((long)z * Mz) + wM * ((long)y * M);
#pragma unroll 4
        for (int x = tx; x < w; x += MP_TX) acc += (double)rp[(long)x * M];
    }
    __shared__ double red[MP_THREADS];
    const int t = ty * MP_TX + tx;
    red[t] = acc;
    __syncthreads();
    for (int st = MP_THREADS / 2; st > 0; st >>= 1) {
        if (t < st) red[t] += red[t + st];
        __syncthreads();
    }
    if (t == 0) mp[o] = (float)red[0] / (float)(w * h * nP);
}

hipError_t launch_mp_map(const float *out, float *mp, int w, int h, int nP, int M, int Mz,
                         hipStream_t s, int nbatch, long ostride, long mstride)
{
    int nentries = 3 * M * M * Mz;
    hipLaunchKernelGGL(k_mp_map, dim3(nentries, nbatch), dim3(MP_TX, MP_TY), 0, s, out, mp, w,
                       h, nP, M, Mz, ostride, mstride);
    return hipGetLastError();
}

// Channels AVG, VAR, TAC2 divided by the MP-map entry of their sub-position (in place).
// Block: 256 pixels of a row; rows (channel, magnified plane, y) over grid.y, images of a
// batch over grid.z (32-bit row arithmetic: 3 * nP * Mz * h * M < 2^31).
__global__ void __launch_bounds__(256) k_mp_correct(float *__restrict__ out,
                                                    const float *__restrict__ mp, int w, int h,
                                                    int nP, int M, int Mz, long ostride,
                                                    long mstride)
{
    out += ostride * blockIdx.z;
    mp += mstride * blockIdx.z;
    const int wM = w * M, hM = h * M, nZ = nP * Mz, MM = M * M;
    const int nrows = 3 * nZ * hM;
    const int xM = blockIdx.x * blockDim.x + threadIdx.x;
    if (xM >= wM) return;
    const int sx = xM % M;
    for (int r = blockIdx.y; r < nrows; r += gridDim.y) {
        const int cz = r / hM, yM = r - cz * hM;     // cz = channel * nZ + zM
        const int ch = cz / nZ, zM = cz - ch * nZ;
        const float d = mp[MM * Mz * ch + MM * (zM % Mz) + M * (yM % M) + sx];
        const long off = (long)r * wM + xM;
        out[off] = out[off] / d;
    }
}

hipError_t launch_mp_correct(float *out, const float *mp, int w, int h, int nP, int M, int Mz,
                             hipStream_t s, int nbatch, long ostride, long mstride)
{
    const int nrows = 3 * nP * Mz * h * M;
    hipLaunchKernelGGL(k_mp_correct, dim3((w * M + 255) / 256, std::min(nrows, 65535), nbatch),
                       dim3(256), 0, s, out, mp, w, h, nP, M, Mz, ostride, mstride);
    return hipGetLastError();
}

__global__ void k_nan_to_zero(const float *__restrict__ src, float *__restrict__ dst, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        float v = src[i];
        dst[i] = (v != v) ? 0.0f : v;
    }
}

hipError_t launch_nan_to_zero_copy(const float *src, float *dst, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_nan_to_zero, dim3(grid_for(n, 256)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

}  // namespace esrrf
