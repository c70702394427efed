// sweep_internal.h -- launch parameters of the parameter-sweep accumulation kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace esrrf {

constexpr int SWEEP_MAX_NF = 8;    // frame counts per sweep
constexpr int SWEEP_MAX_S = 32;    // sensitivities per sweep

struct SweepAccArgs {
    const float *C, *V;    // emitted CGLH and intensity [nframes][M*M][h][w] (FHT: [hM][wM])
    float *out;            // configurations of one magnification, [nf][S][fwhm][4][hM][wM]
    int w, h, M;
    int nframes;           // largest frame count
    int nS, n_fwhm, i_fwhm;
    int nq;
    int int_weighting;
    int fht;               // 1: FHT mode (C, V [frame][hM][wM], Prev = t)
    int nf[SWEEP_MAX_NF];
    float rn[SWEEP_MAX_NF], rn1[SWEEP_MAX_NF];
    float sens[SWEEP_MAX_S];
    int sens_int[SWEEP_MAX_S];
};

hipError_t launch_sweep_acc(const SweepAccArgs &a, hipStream_t s);

}  // namespace esrrf
