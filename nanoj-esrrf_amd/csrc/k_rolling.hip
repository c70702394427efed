// k_rolling.hip -- cumulants of overlapping SR windows (rolling analysis,
// gui/LiveSRRF_optimised_.java:276-384) from per-frame CGLH and intensity.
//
// Without drift a frame's normalised CGLH and interpolated intensity do not depend on the
// SR window it belongs to, so the RGC kernel's emit variant computes them once per frame
// and every window that contains the frame accumulates them.  A block covers 64 magnified
// pixels (in the emit order: phase, Y, X) and one wave per window of the launch: the block
// stages RF frames of the 64 pixels' values in LDS, so HBM sees each frame once per launch,
// and each wave applies the engine's temporal accumulation (cumulate_bic) at the frame's
// index inside its window, so each window's result is the one the engine computes for it.
#include "device_cumulants.h"
#include "esrrf_internal.h"
#include "rolling_internal.h"

namespace esrrf {

constexpr int RF = 64;   // frames staged per pass

// SENS: the sensitivity as an exact integer power 1..4, or 0 for powf (apply_sens)
template <bool INTW, int SENS>
__global__ void __launch_bounds__(64 * ROLL_GROUP) k_roll_acc(const RollAccArgs a)
{
    __shared__ float lC[RF][64], lV[RF][64];
    const long n = (long)a.w * a.h * a.M * a.M;
    const int lane = threadIdx.x & 63;
    const int j = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // window of the wave
    const long o = (long)blockIdx.x * 64 + lane;
    const bool valid = o < n;
    const long wh = (long)a.w * a.h;
    long pix = 0;
    if (valid) {
        const int ph = (int)(o / wh);
        const long r = o - (long)ph * wh;
        const int Y = (int)(r / a.w), X = (int)(r - (long)Y * a.w);
        const int sy = ph / a.M, sx = ph - sy * a.M;
        pix = (long)(Y * a.M + sy) * (a.w * a.M) + (X * a.M + sx);
    }
    const bool wact = j < a.nwin;
    const int win = a.r0 + j;
    const int k0 = a.c0 - win * a.gap;   // index inside the window of the chunk's frame 0
    float *st = a.state + (size_t)(win % a.nslots) * 5 * n;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f, s4 = 0.0f;
    if (wact && valid && k0 > 0) {   // the window started in an earlier chunk
        s0 = st[pix];
        s1 = st[pix + n];
        s2 = st[pix + 2 * n];
        s3 = st[pix + 3 * n];
        s4 = st[pix + 4 * n];
    }
    // frames of the chunk this wave's window covers: [fa, fb)
    const int fa = max(0, -k0), fb = min(a.nframes, a.nF - k0);
    for (int f0 = 0; f0 < a.nframes; f0 += RF) {
        const int nf = min(RF, a.nframes - f0);
        for (int e = threadIdx.x; e < nf * 64; e += blockDim.x) {
            const int fr = e >> 6, p = e & 63;
            const long oo = (long)blockIdx.x * 64 + p;
            const long g = (long)(f0 + fr) * n + oo;
            lC[fr][p] = oo < n ? a.C[g] : 0.0f;
            lV[fr][p] = oo < n ? a.V[g] : 0.0f;
        }
        __syncthreads();
        if (wact) {
            const int lo = max(fa - f0, 0), hi = min(fb - f0, nf);
            for (int fr = lo; fr < hi; fr++)
                cumulate_bic<INTW>(s0, s1, s2, s3, s4, lC[fr][lane], lV[fr][lane],
                                   k0 + f0 + fr == 0, a.sens, SENS, a.rn, a.rn1);
        }
        __syncthreads();
    }
    if (wact && valid) {
        st[pix] = s0;
        st[pix + n] = s1;
        st[pix + 2 * n] = s2;
        st[pix + 3 * n] = s3;
        st[pix + 4 * n] = s4;
    }
}

hipError_t launch_roll_acc(const RollAccArgs &a, hipStream_t s)
{
    if (a.nwin <= 0 || a.nframes <= 0) return hipSuccess;
    if (a.nwin > ROLL_GROUP || a.nslots < 1) return hipErrorInvalidValue;
    const long n = (long)a.w * a.h * a.M * a.M;
    const dim3 grid((unsigned)((n + 63) / 64)), block(64 * ROLL_GROUP);
#define ROLL_LAUNCH(S)                                                                      \
    do {                                                                                    \
        if (a.int_weighting) hipLaunchKernelGGL((k_roll_acc<true, S>), grid, block, 0, s, a); \
        else hipLaunchKernelGGL((k_roll_acc<false, S>), grid, block, 0, s, a);              \
    } while (0)
    switch (a.sens_int) {
    case 1: ROLL_LAUNCH(1); break;
    case 2: ROLL_LAUNCH(2); break;
    case 3: ROLL_LAUNCH(3); break;
    case 4: ROLL_LAUNCH(4); break;
    default: ROLL_LAUNCH(0); break;
    }
#undef ROLL_LAUNCH
    return hipGetLastError();
}

}  // namespace esrrf
