// k_sweep.hip -- cumulants of every (sensitivity, frame count) configuration of a parameter
// sweep from one RGC pass per (magnification, fwhm).
//
// The sensitivity acts on the normalised CGLH of each frame (liveSRRF.cl:358-359) and the
// frame count of a configuration selects the first nf frames of the stack
// (gui/ParametersSweep_.java:800-816), so the per-frame CGLH and intensity that the RGC kernel
// emits serve all of them: each thread streams one pixel's frames once and carries the
// cumulants of every frame count, in the reference's per-configuration order
// (liveSRRF.cl:363-393: t/n and t^2/n added frame by frame, TAC2 with Prev = t/(n-1)).
#include "esrrf_internal.h"
#include "sweep_internal.h"

namespace esrrf {

__device__ __forceinline__ float sweep_sens(float C, float sens, int sens_int)
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

// grid.x: pixels in emit order (phase, Y, X); grid.y: sensitivity index
__global__ void __launch_bounds__(256) k_sweep_acc(const SweepAccArgs a)
{
    const long wh = (long)a.w * a.h;
    const int MM = a.M * a.M;
    const long n = wh * MM;
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const int si = blockIdx.y;
    const float sens = a.sens[si];
    const int sint = a.sens_int[si];
    const int ph = (int)(o / wh);
    const long r = o - (long)ph * wh;
    const int Y = (int)(r / a.w), X = (int)(r - (long)Y * a.w);
    const int sy = ph / a.M, sx = ph - sy * a.M;
    const int wM = a.w * a.M;
    const long whM = wh * MM;
    const long pix = (long)(Y * a.M + sy) * wM + (X * a.M + sx);
    float s0[SWEEP_MAX_NF], s1[SWEEP_MAX_NF], s2[SWEEP_MAX_NF], s3[SWEEP_MAX_NF],
        s4[SWEEP_MAX_NF];
#pragma unroll
    for (int q = 0; q < SWEEP_MAX_NF; q++) s0[q] = s1[q] = s2[q] = s3[q] = s4[q] = 0.0f;
    const bool intw = a.int_weighting != 0;
    for (int f = 0; f < a.nframes; f++) {
        const float C = a.C[(long)f * n + o];
        const float v = a.V[(long)f * n + o];
        const float Cs = sweep_sens(C, sens, sint);
        const float t = intw ? v * Cs : Cs;
        const float e2 = intw ? (t * v) * Cs : Cs * Cs;
#pragma unroll
        for (int q = 0; q < SWEEP_MAX_NF; q++) {
            if (q < a.nq && f < a.nf[q]) {
                const float rn = a.rn[q], rn1 = a.rn1[q];
                if (f == 0) {
                    s0[q] = t * rn;
                    s1[q] = e2 * rn;
                    s2[q] = 0.0f;
                    s3[q] = v * rn;
                } else {
                    s0[q] = s0[q] + t * rn;
                    s1[q] = s1[q] + e2 * rn;
                    s2[q] = s2[q] + (t * s4[q]) * rn1;
                    s3[q] = s3[q] + v * rn;
                }
                s4[q] = t * rn1;   // BIC: Prev = t/(n-1) (liveSRRF.cl:374,377)
                if (f == a.nf[q] - 1) {
                    float *out = a.out + ((long)(q * a.nS + si) * a.n_fwhm + a.i_fwhm) * 4 * whM;
                    out[pix] = s0[q];
                    out[pix + whM] = s1[q];
                    out[pix + 2 * whM] = s2[q];
                    out[pix + 3 * whM] = s3[q];
                }
            }
        }
    }
}

hipError_t launch_sweep_acc(const SweepAccArgs &a, hipStream_t s)
{
    const long n = (long)a.w * a.h * a.M * a.M;
    const dim3 grid((unsigned)((n + 255) / 256), a.nS);
    hipLaunchKernelGGL(k_sweep_acc, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace esrrf
