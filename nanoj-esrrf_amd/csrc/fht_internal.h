// fht_internal.h -- FHT-mode (Fourier-upsampled) launch parameters.
#pragma once
#include <hip/hip_runtime.h>

#include "esrrf_internal.h"

namespace esrrf {

struct FhtPlan {
    int N, log2N, M, w, h;
    int scaleW, scaleH, xROIp, yROIp;   // mirror padding (FourierInterpolation.java:118-172)
    int xROI, yROI;                     // crop of the MN x MN result (:50-54)
    float2 *F;                          // N x N spectrum scratch
    float2 *Y;                          // M^2 x N x N twiddled spectra
    const float2 *tw;                   // N/2 twiddles e^{-2 pi i k / N}
};

struct FhtRgcArgs {
    const float2 *U;
    const float *px;
    float *out, *prev;
    const Tap *taps;
    const int *tapk;
    int W, H;                 // magnified frame
    int frame0, nframes, first_init;
    float rn, rn1, sens;
    int sens_int, int_weighting;
    int R, TW, TH, pitch;     // window half-size, LDS tile extent and pitch
    int ntaps;
    float wsum;
    int tiles_x, tiles_y;
};

struct FhtGenericArgs {
    const float2 *U;
    const float *px;
    const float *drift;
    float *out, *prev;
    int W, H;
    int frame0, nframes, count0, nFrameForSRRF;
    float rn, rn1, sens;
    int sens_int, int_weighting;
    float radius, tSS, tSO;
};

hipError_t fft2_batched(float2 *data, int N, int log2N, int batches, const float2 *tw,
                        hipStream_t s);
hipError_t launch_fht_upsample(const FhtPlan &p, const float *frame, float *out, hipStream_t s);
hipError_t launch_fht_prep(const float *px, float2 *U, int W, int H, int nF, hipStream_t s);
hipError_t launch_fht_grad(const float *px, float *gx, float *gy, int W, int H, hipStream_t s);
size_t fht_rgc_lds_bytes(int TH, int pitch, int ntaps);
hipError_t launch_fht_rgc(const FhtRgcArgs &a, hipStream_t s);
hipError_t launch_fht_generic(const FhtGenericArgs &a, hipStream_t s);

}  // namespace esrrf
