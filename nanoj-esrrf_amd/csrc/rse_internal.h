// rse_internal.h -- buffers and launches of the RSE value/gradient pipeline.
#pragma once
#include <hip/hip_runtime.h>

namespace esrrf {

struct RsePlan {
    int w, h, M;           // raw frame, magnification
    int crop;              // SR pixels cropped on every side (cropSize * M)
    int Wc, Hc;            // cropped size
    long n;                // Wc * Hc
    int mp;                // MP-correct the AVG image
    int K;                 // blur half-width (fixed per problem)
    double *stats;         // [3][M*M] means of AVG, dAVG/dtSS, dAVG/dS per sub-position
    float *Y, *SR, *Dt, *DS;      // cropped ref, SR and its derivatives
    float *T1, *T2, *X, *G1, *G2, *G3;   // blur scratch and blurred images
    float *taps;           // [2][2K+1]
    double *part;          // [rse_max_parts(n)][4]
    double *red;           // [10]: sums, alpha, beta, E, E-gradient dot products
};

hipError_t launch_rse_prep(const RsePlan &p, const float *out4, hipStream_t s);
// value (grad = false) or value and the dot products of the gradient (grad = true)
hipError_t launch_rse_eval(const RsePlan &p, float sigma, bool grad, hipStream_t s);
int rse_max_parts(long n);

}  // namespace esrrf
