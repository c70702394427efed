// k_rolling.hip -- cumulants of overlapping SR windows (rolling analysis,
// gui/LiveSRRF_optimised_.java:276-384) from per-frame CGLH and intensity.
//
// Without drift a frame's normalised CGLH and interpolated intensity do not depend on the
// SR window it belongs to, so the RGC kernel's emit variant computes them once per frame
// and every window that contains the frame accumulates them.  One lane per magnified pixel,
// in the emit order (phase, Y, X), so that the frame values and the window states are read
// and written in consecutive addresses; each lane keeps its pixel's cumulants of every window
// of the launch in registers and walks the chunk's frames once.  A frame's terms t, e2 and
// the TAC2 product t Prev do not depend on the window either (every window has nF frames, so
// the same 1/nF and 1/(nF - 1)), so they are formed once per frame and each window that holds
// the frame adds them with the engine's operations (cumulate_bic in device_cumulants.h): the
// same values, rounded the same way, in the same order.
#include <algorithm>

#include "device_cumulants.h"
#include "esrrf_internal.h"
#include "rolling_internal.h"

namespace esrrf {

constexpr int ROLL_BLOCK = 256;
constexpr int ROLL_PF = 8;   // frames loaded ahead (a multiple of 4)

// SENS: the sensitivity as an exact integer power 1..4, or 0 for powf (apply_sens)
template <bool INTW, int SENS>
__global__ void __launch_bounds__(ROLL_BLOCK) k_roll_acc(const RollAccArgs a)
{
    const long n = (long)a.w * a.h * a.M * a.M;
    const long o = (long)blockIdx.x * ROLL_BLOCK + threadIdx.x;
    const bool valid = o < n;
    const long ol = valid ? o : n - 1;   // in-bounds address for the loads of idle lanes
    // window r0 + j holds chunk frame f at index k0 - j gap + f
    const int k0 = a.c0 - a.r0 * a.gap;
    // A window's first frame sets AVG = t/nF, E2 = e2/nF, TAC2 = 0, INT = v/nF
    // (liveSRRF.cl:364-368).  fmaf(x, r, -0) rounds x r exactly as x * r does (signed zeros
    // included), so a window that starts in this chunk begins at -0 and takes the same fmaf
    // as its later frames; its TAC2 keeps 0 at its first frame.
    float s[ROLL_GROUP][4];
#pragma unroll
    for (int j = 0; j < ROLL_GROUP; j++) {
        s[j][0] = s[j][1] = s[j][3] = -0.0f;
        s[j][2] = 0.0f;
        if (j < a.nwin && k0 - j * a.gap > 0) {   // the window started in an earlier chunk
            const float *st = a.state + (size_t)((a.r0 + j) % a.nslots) * 5 * n;
            s[j][0] = st[ol];
            s[j][1] = st[ol + n];
            s[j][2] = st[ol + 2 * n];
            s[j][3] = st[ol + 3 * n];
        }
    }
    // Prev of the frame before the chunk (the same in every window that holds it)
    float prv = 0.0f;
    if (k0 > 0) prv = a.state[(size_t)(a.r0 % a.nslots) * 5 * n + 4 * n + ol];
    // frames of the chunk some window of the launch holds: [fa, fb)
    const int fa = max(0, -k0);
    const int fb = min(a.nframes, a.nF - (k0 - (a.nwin - 1) * a.gap));
    // window j holds chunk frames [j gap - k0, j gap - k0 + nF): windows [0, nst) have started
    // and windows [0, nen) have ended by the frame in hand (scalar counters, kept per frame)
    int nst = 0, nen = 0;
    while (nst < a.nwin && nst * a.gap - k0 < fa) nst++;
    const float *Cb = a.C + ol, *Vb = a.V + ol;
    // every load is issued (frame indices clamped to the last one), so none sits under a branch
    float cC[ROLL_PF], cV[ROLL_PF];
#pragma unroll
    for (int q = 0; q < ROLL_PF; q++) {
        const long fq = min(fa + q, fb - 1);
        cC[q] = Cb[fq * n];
        cV[q] = Vb[fq * n];
    }
    for (int f0 = fa; f0 < fb; f0 += ROLL_PF) {
        float nC[ROLL_PF], nV[ROLL_PF];
#pragma unroll
        for (int q = 0; q < ROLL_PF; q++) {   // the next frames' loads, in flight meanwhile
            const long fq = min(f0 + ROLL_PF + q, fb - 1);
            nC[q] = Cb[fq * n];
            nV[q] = Vb[fq * n];
        }
        // groups of four frames: a window that holds all four at an index > 0 takes them in one
        // branch; windows that start or end inside the group take them frame by frame
#pragma unroll
        for (int g = 0; g < ROLL_PF; g += 4) {
            unsigned fm[4], lm[4];   // per frame: windows at index 0, at an index > 0
            float t[4], e2[4], tp[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int f = f0 + g + q;
                fm[q] = lm[q] = 0u;
                if (f < fb) {
                    const int st0 = nst;
                    while (nst < a.nwin && nst * a.gap - k0 <= f) nst++;
                    while (nen < a.nwin && nen * a.gap - k0 + a.nF <= f) nen++;
                    fm[q] = ((1u << nst) - 1u) & ~((1u << st0) - 1u);
                    lm[q] = ((1u << st0) - 1u) & ~((1u << nen) - 1u);
                }
                bic_terms<INTW>(cC[g + q], cV[g + q], a.sens, SENS, t[q], e2[q]);
                tp[q] = __fmul_rn(t[q], prv);
                if (f < fb) prv = __fmul_rn(t[q], a.rn1);
            }
            const unsigned all4 = lm[0] & lm[1] & lm[2] & lm[3];
#pragma unroll
            for (int j = 0; j < ROLL_GROUP; j++)
                if (all4 & (1u << j)) {
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        s[j][0] = fmaf(t[q], a.rn, s[j][0]);
                        s[j][1] = fmaf(e2[q], a.rn, s[j][1]);
                        s[j][2] = fmaf(tp[q], a.rn1, s[j][2]);
                        s[j][3] = fmaf(cV[g + q], a.rn, s[j][3]);
                    }
                }
            unsigned rest = 0u;
#pragma unroll
            for (int q = 0; q < 4; q++) rest |= fm[q] | (lm[q] & ~all4);
            if (rest) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const unsigned lq = lm[q] & ~all4;
#pragma unroll
                    for (int j = 0; j < ROLL_GROUP; j++) {
                        if (lq & (1u << j)) {
                            s[j][0] = fmaf(t[q], a.rn, s[j][0]);
                            s[j][1] = fmaf(e2[q], a.rn, s[j][1]);
                            s[j][2] = fmaf(tp[q], a.rn1, s[j][2]);
                            s[j][3] = fmaf(cV[g + q], a.rn, s[j][3]);
                        }
                        if (fm[q] & (1u << j)) {   // TAC2 stays 0
                            s[j][0] = fmaf(t[q], a.rn, s[j][0]);
                            s[j][1] = fmaf(e2[q], a.rn, s[j][1]);
                            s[j][3] = fmaf(cV[g + q], a.rn, s[j][3]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < ROLL_PF; q++) {
            cC[q] = nC[q];
            cV[q] = nV[q];
        }
    }
    if (!valid) return;
#pragma unroll
    for (int j = 0; j < ROLL_GROUP; j++) {
        if (j < a.nwin) {
            float *st = a.state + (size_t)((a.r0 + j) % a.nslots) * 5 * n;
            st[o] = s[j][0];
            st[o + n] = s[j][1];
            st[o + 2 * n] = s[j][2];
            st[o + 3 * n] = s[j][3];
            st[o + 4 * n] = prv;   // Prev after the chunk: read by the next chunk's launch
        }
    }
}

hipError_t launch_roll_acc(const RollAccArgs &a, hipStream_t s)
{
    if (a.nwin <= 0 || a.nframes <= 0) return hipSuccess;
    if (a.nwin > ROLL_GROUP || a.nslots < 1) return hipErrorInvalidValue;
    const long n = (long)a.w * a.h * a.M * a.M;
    const dim3 grid((unsigned)((n + ROLL_BLOCK - 1) / ROLL_BLOCK)), block(ROLL_BLOCK);
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

// Channels 0..3 of a window from the emit order [c][M*M][h][w] to the image order
// [c][h M][w M], one thread per output pixel (rows over grid.y, channels over grid.z).
__global__ void __launch_bounds__(256) k_roll_to_image(const float *__restrict__ src,
                                                       float *__restrict__ dst, int w, int h,
                                                       int M)
{
    const int wM = w * M, hM = h * M;
    const long n = (long)wM * hM;
    const int xM = blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.z;
    if (xM >= wM) return;
    const int X = xM / M, sx = xM - X * M;
    for (int yM = blockIdx.y; yM < hM; yM += gridDim.y) {
        const int Y = yM / M, sy = yM - Y * M;
        const long i = ((long)(sy * M + sx) * h + Y) * w + X;
        dst[c * n + (long)yM * wM + xM] = src[c * n + i];
    }
}

hipError_t launch_roll_to_image(const float *src, float *dst, int w, int h, int M, hipStream_t s)
{
    const dim3 grid((unsigned)((w * M + 255) / 256), (unsigned)std::min(h * M, 65535), 4);
    hipLaunchKernelGGL(k_roll_to_image, grid, dim3(256), 0, s, src, dst, w, h, M);
    return hipGetLastError();
}

}  // namespace esrrf
