// k_squirrel.hip -- SQUIRREL error map of a (reference, SR) image pair
// (ErrorMapLiveSRRF.java:88-190); the blur, the alpha/beta fit and the residual sums are the
// RSE kernels of k_rse.hip.
//   k_sq_crop    crop cropSize*M pixels per side of ref and SR (:88-99); NaN -> 0
//   k_sq_maps    RSC = alpha blur(SR) + beta (= blur(alpha SR + beta), the blur being
//                normalised with clamp-to-edge borders), the error map |ref - RSC| (:180-188)
//                and per-block sums of RSC, ref, RSC^2, ref^2, RSC ref, (RSC - ref)^2
//   k_sq_finish  RSE = sqrt(mean((RSC - ref)^2)) and the Pearson correlation of RSC and ref
//                (globalRMSE, globalPPMCC, :175-176)
#include "rse_internal.h"

namespace esrrf {

constexpr int SB = 256;

__device__ __forceinline__ double sq_block_sum(double v, double *sh)
{
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = SB / 2; s > 0; s >>= 1) {
        if (tid < s) sh[tid] += sh[tid + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ float finite_or_zero(float v) { return v == v ? v : 0.0f; }

__global__ void k_sq_crop(const float *__restrict__ ref, const float *__restrict__ sr, int W,
                          int crop, int Wc, int Hc, float *__restrict__ Y, float *__restrict__ SR)
{
    const long n = (long)Wc * Hc;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const int y = (int)(i / Wc), x = (int)(i - (long)y * Wc);
        const long g = (long)(y + crop) * W + (x + crop);
        Y[i] = finite_or_zero(ref[g]);
        SR[i] = finite_or_zero(sr[g]);
    }
}

__global__ void __launch_bounds__(SB) k_sq_maps(const float *__restrict__ Y,
                                                const float *__restrict__ X,
                                                const double *__restrict__ red, long n,
                                                float *__restrict__ emap,
                                                float *__restrict__ rsc,
                                                double *__restrict__ part)
{
    __shared__ double sh[SB];
    const float a = (float)red[4], b = (float)red[5];
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (long i = (long)blockIdx.x * SB + threadIdx.x; i < n; i += (long)gridDim.x * SB) {
        const float c = fmaf(a, X[i], b);
        const float y = Y[i];
        if (rsc) rsc[i] = c;
        if (emap) emap[i] = fabsf(y - c);
        const double cd = c, yd = y, dd = (double)c - (double)y;
        s[0] += cd;
        s[1] += yd;
        s[2] += cd * cd;
        s[3] += yd * yd;
        s[4] += cd * yd;
        s[5] += dd * dd;
    }
    for (int k = 0; k < 6; k++) s[k] = sq_block_sum(s[k], sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) part[6 * blockIdx.x + k] = s[k];
}

// out[0] = RSE, out[1] = RSP, out[2..3] = alpha, beta
__global__ void __launch_bounds__(SB) k_sq_finish(const double *__restrict__ part, int nparts,
                                                  long n, const double *__restrict__ red,
                                                  double *__restrict__ out)
{
    __shared__ double sh[SB];
    double t[6] = {0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nparts; b += SB)
        for (int k = 0; k < 6; k++) t[k] += part[6 * b + k];
    for (int k = 0; k < 6; k++) t[k] = sq_block_sum(t[k], sh);
    if (threadIdx.x == 0) {
        const double N = (double)n;
        const double mx = t[0] / N, my = t[1] / N;
        const double vx = t[2] / N - mx * mx, vy = t[3] / N - my * my;
        const double cxy = t[4] / N - mx * my;
        out[0] = sqrt(fmax(t[5], 0.0) / N);
        out[1] = (vx > 0.0 && vy > 0.0) ? cxy / sqrt(vx * vy) : 0.0;
        out[2] = (float)red[4];
        out[3] = (float)red[5];
    }
}

static unsigned sq_grid(long n)
{
    long g = (n + SB - 1) / SB;
    return (unsigned)(g > 2048 ? 2048 : (g < 1 ? 1 : g));
}

int squirrel_max_parts(long n) { return (int)sq_grid(n); }

hipError_t launch_squirrel_crop(const float *ref, const float *sr, int W, int crop,
                                const RsePlan &p, hipStream_t s)
{
    hipLaunchKernelGGL(k_sq_crop, dim3(sq_grid(p.n)), dim3(SB), 0, s, ref, sr, W, crop, p.Wc,
                       p.Hc, p.Y, p.SR);
    const hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_rse_yy(p, s);   // sum y^2 of the new reference
}

hipError_t launch_squirrel_maps(const RsePlan &p, float *emap, float *rsc, double *part,
                                double *out, hipStream_t s)
{
    const unsigned g = sq_grid(p.n);
    hipLaunchKernelGGL(k_sq_maps, dim3(g), dim3(SB), 0, s, p.Y, p.X, p.red, p.n, emap, rsc, part);
    hipLaunchKernelGGL(k_sq_finish, dim3(1), dim3(SB), 0, s, part, (int)g, p.n, p.red, out);
    return hipGetLastError();
}

}  // namespace esrrf
