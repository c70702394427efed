// fht_internal.h -- FHT-mode (Fourier-upsampled) launch parameters.
#pragma once
#include <hip/hip_runtime.h>

#include "esrrf_internal.h"

namespace esrrf {

struct FhtPlan {
    int N, log2N, M, w, h;
    int scaleW, scaleH, xROIp, yROIp;   // mirror padding (FourierInterpolation.java:118-172)
    int xROI, yROI;                     // crop of the MN x MN result (:50-54)
    int chunk;                          // frames per batch of the scratch spectra
    int scatter;                        // 1: last pass stores pixel by pixel (k_fht_inv_cols)
    int xcd;                            // 1: column passes take tiles in XCD-grouped order
    int no_prefetch;                    // 1: last pass without the register prefetch
    float2 *F;                          // [chunk][N][N] spectra
    float2 *Y;                          // [chunk][M][N][N] row-transformed twiddled spectra
    const float2 *tw;                   // N twiddles e^{-2 pi i k / N}
    const float2 *twMN;                 // M*N twiddles e^{-2 pi i j / (M N)}
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
    int Rhi;                  // window reach on the high side ((int)(radius + 1))
    int pair;                 // 1: two frames and two pixels per lane where LDS allows
    int ntaps;
    float wsum;
    int tiles_x, tiles_y;
    // non-null: emit each pixel's normalised CGLH (before the sensitivity power) per frame,
    // [frame][H][W], instead of accumulating (the parameter sweep)
    float *emitC;
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

int fht_group(int N);
hipError_t launch_fht_upsample(const FhtPlan &p, const float *frames, float *out, int nf,
                               hipStream_t s);
hipError_t launch_fht_prep(const float *px, float2 *U, int W, int H, int nF, hipStream_t s);
hipError_t launch_fht_grad(const float *px, float *gx, float *gy, int W, int H, hipStream_t s);
size_t fht_rgc_lds_bytes(int TH, int pitch, int ntaps);
hipError_t launch_fht_rgc(const FhtRgcArgs &a, hipStream_t s, const char **kname = nullptr);
hipError_t launch_fht_generic(const FhtGenericArgs &a, hipStream_t s);

}  // namespace esrrf
