// device_cumulants.h -- the sensitivity power and the BIC temporal accumulation of one pixel
// (liveSRRF.cl:358-393), shared by the RGC kernels and the kernels that accumulate emitted
// per-frame values, so that both round identically.
#pragma once
#include <hip/hip_runtime.h>

namespace esrrf {

__device__ __forceinline__ float apply_sens(float C, float sens, int sens_int)
{
    if (!(C >= 0.0f)) return 0.0f;   // negative or NaN -> 0 (liveSRRF.cl:358-359)
    switch (sens_int) {
    case 1: return C;
    case 2: return C * C;
    case 3: return C * C * C;
    case 4: { const float c2 = C * C; return c2 * c2; }
    default: return powf(C, sens);
    }
}

// One frame of the temporal accumulation (liveSRRF.cl:363-393) from the normalised CGLH C0
// and the interpolated intensity v: t = v C^S (C^S without intensity weighting); the first
// frame of an SR image initialises; Prev holds t/(n-1) (the BIC TAC2 quirk, :374,377).
// A frame's accumulation terms from the normalised CGLH C0 and the intensity v: t = v C^S and
// e2 = v C^S v C^S (C^S and C^S C^S without intensity weighting) (liveSRRF.cl:357,365-366,379-380).
template <bool INTW>
__device__ __forceinline__ void bic_terms(float C0, float v, float sens, int sens_int, float &t,
                                          float &e2)
{
    const float C = apply_sens(C0, sens, sens_int);
    t = INTW ? __fmul_rn(v, C) : C;
    e2 = INTW ? __fmul_rn(__fmul_rn(t, v), C) : __fmul_rn(C, C);
}

template <bool INTW>
__device__ __forceinline__ void cumulate_bic(float &avg, float &e2s, float &tac, float &in,
                                             float &prv, float C0, float v, bool first,
                                             float sens, int sens_int, float rn, float rn1)
{
    float t, e2;
    bic_terms<INTW>(C0, v, sens, sens_int, t, e2);
    if (first) {
        avg = __fmul_rn(t, rn);
        e2s = __fmul_rn(e2, rn);
        tac = 0.0f;
        in = __fmul_rn(v, rn);
    } else {
        avg = fmaf(t, rn, avg);
        e2s = fmaf(e2, rn, e2s);
        tac = fmaf(__fmul_rn(t, prv), rn1, tac);
        in = fmaf(v, rn, in);
    }
    prv = __fmul_rn(t, rn1);
}

}  // namespace esrrf
