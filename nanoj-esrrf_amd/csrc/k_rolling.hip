// k_rolling.hip -- cumulants of overlapping SR windows (rolling analysis,
// gui/LiveSRRF_optimised_.java:276-384) from per-frame CGLH and intensity.
//
// Without drift a frame's normalised CGLH and interpolated intensity do not depend on the
// SR window it belongs to, so the RGC kernel's emit variant computes them once per frame
// and every window that contains the frame accumulates them.  A thread owns one magnified
// pixel (in the emit order: phase, Y, X) and up to ROLL_GROUP windows; it streams the chunk's
// frames once and applies, for each window, the engine's temporal accumulation
// (cumulate_bic) at the frame's index inside that window, so each window's result is the one
// the engine computes for it.
#include "device_cumulants.h"
#include "esrrf_internal.h"
#include "rolling_internal.h"

namespace esrrf {

template <bool INTW>
__global__ void __launch_bounds__(256) k_roll_acc(const RollAccArgs a)
{
    const long wh = (long)a.w * a.h;
    const int MM = a.M * a.M;
    const long n = wh * MM;
    const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const int ph = (int)(o / wh);
    const long r = o - (long)ph * wh;
    const int Y = (int)(r / a.w), X = (int)(r - (long)Y * a.w);
    const int sy = ph / a.M, sx = ph - sy * a.M;
    const long pix = (long)(Y * a.M + sy) * (a.w * a.M) + (X * a.M + sx);
    const int g0 = blockIdx.y * ROLL_GROUP;
    float s0[ROLL_GROUP], s1[ROLL_GROUP], s2[ROLL_GROUP], s3[ROLL_GROUP], s4[ROLL_GROUP];
#pragma unroll
    for (int j = 0; j < ROLL_GROUP; j++) {
        s0[j] = s1[j] = s2[j] = s3[j] = s4[j] = 0.0f;
        const int wj = g0 + j;
        if (wj < a.nwin && a.k0[wj] > 0) {   // the window started in an earlier chunk
            const float *st = a.state[wj];
            s0[j] = st[pix];
            s1[j] = st[pix + n];
            s2[j] = st[pix + 2 * n];
            s3[j] = st[pix + 3 * n];
            s4[j] = st[pix + 4 * n];
        }
    }
    for (int f = 0; f < a.nframes; f++) {
        const float C0 = a.C[(long)f * n + o];
        const float v = a.V[(long)f * n + o];
#pragma unroll
        for (int j = 0; j < ROLL_GROUP; j++) {
            const int wj = g0 + j;
            if (wj >= a.nwin) continue;
            const int k = a.k0[wj] + f;   // frame index inside window wj
            if (k < 0 || k >= a.nF) continue;
            cumulate_bic<INTW>(s0[j], s1[j], s2[j], s3[j], s4[j], C0, v, k == 0, a.sens,
                               a.sens_int, a.rn, a.rn1);
        }
    }
#pragma unroll
    for (int j = 0; j < ROLL_GROUP; j++) {
        const int wj = g0 + j;
        if (wj >= a.nwin) continue;
        float *st = a.state[wj];
        st[pix] = s0[j];
        st[pix + n] = s1[j];
        st[pix + 2 * n] = s2[j];
        st[pix + 3 * n] = s3[j];
        st[pix + 4 * n] = s4[j];
    }
}

hipError_t launch_roll_acc(const RollAccArgs &a, hipStream_t s)
{
    if (a.nwin <= 0 || a.nframes <= 0) return hipSuccess;
    const long n = (long)a.w * a.h * a.M * a.M;
    const dim3 grid((unsigned)((n + 255) / 256), (unsigned)((a.nwin + ROLL_GROUP - 1) / ROLL_GROUP));
    if (a.int_weighting)
        hipLaunchKernelGGL(k_roll_acc<true>, grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_roll_acc<false>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace esrrf
