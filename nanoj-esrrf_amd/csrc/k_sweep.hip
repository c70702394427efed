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

template <bool POW>
__device__ __forceinline__ float sweep_sens(float C, float sens, int sens_int)
{
    if (!(C >= 0.0f)) return 0.0f;
    if (!POW) {   // integer sensitivities 1..4 only
        const float c2 = C * C;
        return sens_int == 1 ? C : sens_int == 2 ? c2 : sens_int == 3 ? c2 * C : c2 * c2;
    }
    switch (sens_int) {
    case 1: return C;
    case 2: return C * C;
    case 3: return C * C * C;
    case 4: { const float c2 = C * C; return c2 * c2; }
    default: return powf(C, sens);
    }
}

// C^S of a thread's SG sensitivities for one frame.  Integer sensitivities 1..4 (POW false):
// the four powers apply_sens forms (C, C^2, C^3 = C^2 C, C^4 = C^2 C^2 of C, negative or NaN
// -> 0, liveSRRF.cl:358-359) are formed once and each sensitivity takes its power's bits
// through masks set once per thread (all ones for its power, zero for the others): no branch
// in the frame loop, the same values as sweep_sens.
template <int SG, bool POW>
struct SensPow {
    unsigned m[SG][4];
    float sens[SG];
    int sint[SG];
    __device__ __forceinline__ void init(const float *s, const int *si, int si0, int nS)
    {
#pragma unroll
        for (int g = 0; g < SG; g++) {
            const int k = min(si0 + g, nS - 1);
            sens[g] = s[k];
            sint[g] = si[k];
#pragma unroll
            for (int e = 0; e < 4; e++) m[g][e] = sint[g] == e + 1 ? 0xffffffffu : 0u;
        }
    }
    __device__ __forceinline__ void apply(float C, float (&Cs)[SG]) const
    {
        if (POW) {
#pragma unroll
            for (int g = 0; g < SG; g++) Cs[g] = sweep_sens<true>(C, sens[g], sint[g]);
            return;
        }
        const float c1 = C >= 0.0f ? C : 0.0f;
        const float c2 = c1 * c1;
        const unsigned b1 = __float_as_uint(c1), b2 = __float_as_uint(c2),
                       b3 = __float_as_uint(c2 * c1), b4 = __float_as_uint(c2 * c2);
#pragma unroll
        for (int g = 0; g < SG; g++)
            Cs[g] = __uint_as_float((b1 & m[g][0]) | (b2 & m[g][1]) | (b3 & m[g][2]) |
                                    (b4 & m[g][3]));
    }
};

// grid.x: pixels in emit order (phase, Y, X); grid.y: groups of SG sensitivities.  A thread
// reads each frame's CGLH and intensity of its pixel once and updates the cumulants of the
// NQ frame counts of each of its SG sensitivities (the frames are read once per group, not
// once per sensitivity).
template <int NQ, int SG, bool POW>
__global__ void __launch_bounds__(256) k_sweep_acc(const SweepAccArgs a)
{
    const long wh = (long)a.w * a.h;
    const int MM = a.M * a.M;
    const long n = wh * MM;
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const int si0 = blockIdx.y * SG;
    SensPow<SG, POW> sp;
    sp.init(a.sens, a.sens_int, si0, a.nS);
    const int ph = (int)(o / wh);
    const long r = o - (long)ph * wh;
    const int Y = (int)(r / a.w), X = (int)(r - (long)Y * a.w);
    const int sy = ph / a.M, sx = ph - sy * a.M;
    const int wM = a.w * a.M;
    const long whM = wh * MM;
    const long pix = (long)(Y * a.M + sy) * wM + (X * a.M + sx);
    // per (sensitivity, frame count): AVG, E2 and TAC2 sums; the INT sum does not depend on
    // the sensitivity, and Prev = t/(n-1) is recomputed from the previous frame's t (the
    // same rounded product the reference keeps)
    float s0[SG][NQ], s1[SG][NQ], s2[SG][NQ], s3[NQ], tp[SG];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        s3[q] = 0.0f;
#pragma unroll
        for (int g = 0; g < SG; g++) s0[g][q] = s1[g][q] = s2[g][q] = 0.0f;
    }
#pragma unroll
    for (int g = 0; g < SG; g++) tp[g] = 0.0f;
    const bool intw = a.int_weighting != 0;
    if (a.nframes <= 0) return;
    {   // frame 0 initialises every configuration (each holds at least one frame)
        const float C = a.C[o];
        const float v = a.V[o];
#pragma unroll
        for (int q = 0; q < NQ; q++) s3[q] = v * a.rn[q];
        float Csg[SG];
        sp.apply(C, Csg);
#pragma unroll
        for (int g = 0; g < SG; g++) {
            const float Cs = Csg[g];
            const float t = intw ? v * Cs : Cs;
            const float e2 = intw ? (t * v) * Cs : Cs * Cs;
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                s0[g][q] = t * a.rn[q];
                s1[g][q] = e2 * a.rn[q];
                s2[g][q] = 0.0f;
            }
            tp[g] = t;
        }
    }
    // frames 1 .. nf[q] - 1 of configuration q.  Frame 0 outside the loop and 1/n, 1/(n-1)
    // set to 0 (a scalar select) for a configuration past its last frame keep the loop body
    // free of branches: such a configuration adds products t * 0 = +-0, which leave its sums'
    // values unchanged.
#pragma unroll 4
    for (int f = 1; f < a.nframes; f++) {
        const float C = a.C[(long)f * n + o];
        const float v = a.V[(long)f * n + o];
        float rnq[NQ], rn1q[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const bool on = q < a.nq && f < a.nf[q];
            rnq[q] = on ? a.rn[q] : 0.0f;
            rn1q[q] = on ? a.rn1[q] : 0.0f;
            s3[q] = s3[q] + v * rnq[q];
        }
        float Csg[SG];
        sp.apply(C, Csg);
#pragma unroll
        for (int g = 0; g < SG; g++) {
            const float Cs = Csg[g];
            const float t = intw ? v * Cs : Cs;
            const float e2 = intw ? (t * v) * Cs : Cs * Cs;
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                s0[g][q] = s0[g][q] + t * rnq[q];
                s1[g][q] = s1[g][q] + e2 * rnq[q];
                // BIC: Prev = t/(n-1) (liveSRRF.cl:374,377)
                s2[g][q] = s2[g][q] + (t * (tp[g] * rn1q[q])) * rn1q[q];
            }
            tp[g] = t;
        }
    }
    // the sums of frame count q stop changing after its last frame nf[q] - 1
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        if (q >= a.nq) continue;
#pragma unroll
        for (int g = 0; g < SG; g++) {
            if (si0 + g >= a.nS) continue;
            float *out = a.out + ((long)(q * a.nS + si0 + g) * a.n_fwhm + a.i_fwhm) * 4 * whM;
            out[pix] = s0[g][q];
            out[pix + whM] = s1[g][q];
            out[pix + 2 * whM] = s2[g][q];
            out[pix + 3 * whM] = s3[q];
        }
    }
}

// FHT mode (liveSRRF_noInterpolation.cl:358-390): one thread per magnified pixel o; the
// emitted CGLH is [frame][hM][wM] and the intensity is the upsampled frame itself (a.V, the
// same layout); TAC2 keeps Prev = t (single 1/(n-1)).  The operations are those of the
// FHT engine's per-frame accumulation, spelled out, so every configuration equals the
// per-configuration engine bit for bit.
template <int NQ, int SG, bool POW>
__global__ void __launch_bounds__(256) k_sweep_acc_fht(const SweepAccArgs a)
{
    const long n = (long)a.w * a.M * a.h * a.M;
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const int si0 = blockIdx.y * SG;
    SensPow<SG, POW> sp;
    sp.init(a.sens, a.sens_int, si0, a.nS);
    float s0[SG][NQ], s1[SG][NQ], s2[SG][NQ], s3[NQ], tp[SG];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        s3[q] = 0.0f;
#pragma unroll
        for (int g = 0; g < SG; g++) s0[g][q] = s1[g][q] = s2[g][q] = 0.0f;
    }
#pragma unroll
    for (int g = 0; g < SG; g++) tp[g] = 0.0f;
    const bool intw = a.int_weighting != 0;
#pragma unroll 4
    for (int f = 0; f < a.nframes; f++) {
        const float C = a.C[(long)f * n + o];
        const float v = a.V[(long)f * n + o];
#pragma unroll
        for (int q = 0; q < NQ; q++)
            if (q < a.nq && f < a.nf[q])
                s3[q] = f == 0 ? __fmul_rn(v, a.rn[q]) : fmaf(v, a.rn[q], s3[q]);
        float Csg[SG];
        sp.apply(C, Csg);
#pragma unroll
        for (int g = 0; g < SG; g++) {
            const float Cs = Csg[g];
            const float t = intw ? __fmul_rn(v, Cs) : Cs;
            const float e2 = intw ? __fmul_rn(__fmul_rn(t, v), Cs) : __fmul_rn(Cs, Cs);
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                if (q < a.nq && f < a.nf[q]) {
                    const float rn = a.rn[q], rn1 = a.rn1[q];
                    if (f == 0) {
                        s0[g][q] = __fmul_rn(t, rn);
                        s1[g][q] = __fmul_rn(e2, rn);
                        s2[g][q] = 0.0f;
                    } else {
                        s0[g][q] = fmaf(t, rn, s0[g][q]);
                        s1[g][q] = fmaf(e2, rn, s1[g][q]);
                        s2[g][q] = fmaf(__fmul_rn(t, tp[g]), rn1, s2[g][q]);
                    }
                }
            }
            tp[g] = t;
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        if (q >= a.nq) continue;
#pragma unroll
        for (int g = 0; g < SG; g++) {
            if (si0 + g >= a.nS) continue;
            float *out = a.out + ((long)(q * a.nS + si0 + g) * a.n_fwhm + a.i_fwhm) * 4 * n;
            out[o] = s0[g][q];
            out[o + n] = s1[g][q];
            out[o + 2 * n] = s2[g][q];
            out[o + 3 * n] = s3[q];
        }
    }
}

template <int NQ, int SG>
static hipError_t launch_acc(const SweepAccArgs &a, hipStream_t s)
{
    const long n = (long)a.w * a.h * a.M * a.M;
    const dim3 grid((unsigned)((n + 255) / 256), (unsigned)((a.nS + SG - 1) / SG));
    bool pow = false;
    for (int i = 0; i < a.nS; i++) pow = pow || a.sens_int[i] < 1 || a.sens_int[i] > 4;
    if (a.fht) {
        if (pow)
            hipLaunchKernelGGL((k_sweep_acc_fht<NQ, SG, true>), grid, dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_sweep_acc_fht<NQ, SG, false>), grid, dim3(256), 0, s, a);
    } else if (pow) {
        hipLaunchKernelGGL((k_sweep_acc<NQ, SG, true>), grid, dim3(256), 0, s, a);
    } else {
        hipLaunchKernelGGL((k_sweep_acc<NQ, SG, false>), grid, dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

// about 60 cumulant registers per thread: 4 sensitivities for up to 3 frame counts, 2 for 4,
// 1 for up to SWEEP_MAX_NF
hipError_t launch_sweep_acc(const SweepAccArgs &a, hipStream_t s)
{
    if (a.nq <= 1) return launch_acc<1, 4>(a, s);
    if (a.nq <= 2) return launch_acc<2, 4>(a, s);
    if (a.nq <= 3) return launch_acc<3, 4>(a, s);
    if (a.nq <= 4) return launch_acc<4, 2>(a, s);
    return launch_acc<SWEEP_MAX_NF, 1>(a, s);
}

}  // namespace esrrf
