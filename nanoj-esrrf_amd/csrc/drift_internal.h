// drift_internal.h -- drift estimation against a reference frame: plan and launches.
#pragma once
#include <hip/hip_runtime.h>

namespace esrrf {

struct DriftPlan {
    int w, h;                 // frames
    int N, log2N;             // transform size (power of two >= max(w, h))
    int max_batch;            // frames per batch (buf holds max_batch transforms)
    const float2 *tw;         // [N] e^{-2 pi i t / N}
    float2 *R;                // [N][N] spectrum of the reference frame
    float2 *buf;              // [max_batch][N][N]
    double *sums;             // [max_batch] frame sums, then [max_batch][DRIFT_SEG] partials
    float *pkv;               // [max_batch][DRIFT_SEG] segment maxima of the correlation
    int *pki;                 // [max_batch][DRIFT_SEG] their first indices
};

// blocks per frame of the sum and peak searches (each takes a contiguous segment)
constexpr int DRIFT_SEG = 16;

// spectrum of the mean-free, zero-padded reference frame into d.R
hipError_t launch_drift_reference(const DriftPlan &d, const float *ref, hipStream_t s);
// drift of `batch` consecutive frames against d.R into sx, sy (device, batch entries each)
hipError_t launch_drift_batch(const DriftPlan &d, const float *frames, int batch, float *sx,
                              float *sy, hipStream_t s);

}  // namespace esrrf
