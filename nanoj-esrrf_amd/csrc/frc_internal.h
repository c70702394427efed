// frc_internal.h -- Fourier ring correlation of two images: plan and launches.
#pragma once
#include <hip/hip_runtime.h>

namespace esrrf {

constexpr int FRC_MAX_N = 4096;    // largest transform (ring bins of one block fit in LDS)

struct FrcPlan {
    int W, H;                 // input images
    int S, x0, y0;            // centred square crop
    int N, log2N;             // transform size (power of two >= S)
    const float *win;         // [S] Tukey window
    const float2 *tw;         // [N] e^{-2 pi i t / N}
    float2 *buf;              // [2][N][N] spectra of the two images
    double *part;             // [nparts][3][N/2] ring sums per block
    double *rings;            // [3][N/2]: sum Re(F1 conj F2), sum |F1|^2, sum |F2|^2
    int nparts;
};

hipError_t launch_frc(const FrcPlan &p, const float *a, const float *b, hipStream_t s);

// Resolution N / r* (pixels) at the first crossing r* of 1/7 after ring 0 of the curve
// FRC(r) = num / sqrt(d1 d2) from the ring sums [3][N/2] (linear between rings; NaN when
// the curve never crosses); the curve into `curve` (N/2 values) when given.
double frc_resolution(const double *rings, int N, float *curve);

// Forward 2-D FFT, in place, of `batch` N x N complex images at p.buf (p.N, p.log2N, p.tw).
hipError_t launch_fft2d(const FrcPlan &p, int batch, hipStream_t s);

}  // namespace esrrf
