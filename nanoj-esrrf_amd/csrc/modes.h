// modes.h -- FHT (Fourier-upsampled) and 3-D (multi-focus) pipelines of the engine.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/esrrf.h"

namespace esrrf {

struct ModeState {
    int mode = 0;                 // 1 = FHT, 2 = 3-D
    std::string err;
    double prof_up_ms = 0.0;      // upsampling kernel time (FHT)
    long prof_up_n = 0;
    long long active = 0;         // gated RGC samples per frame (geometry only)
    // FHT
    int N = 0;                    // padded square size of the forward transform
    float2 *d_spec = nullptr;     // spectra / FFT scratch
    size_t spec_bytes = 0;
    float *d_g = nullptr;         // RobX gradients of the load [2][nF][hM][wM]
    void *taps = nullptr;         // device tap table (FHT / 3-D)
    int ntaps = 0;
    float tap_wsum = 0.0f;
    int tap_pp = 0, tap_ph = 0, tap_tw = 0, tap_th = 0, tap_r = 0, tap_rhi = 0;
    // 3-D
    float *d_st = nullptr;        // [3][nP] theta(rad), shiftX, shiftY (LiveSRRF_CL.java:440-445)
    double *d_roi = nullptr, *d_ic = nullptr;
    float bg = 0.0f;
    float *d_gi = nullptr;        // interpolated gradients [3][nF][nP][2h][2w]
    int nF = 0;
};

struct ModeCtx {
    hipStream_t s;
    int w, h, W, H, nP, M, wM, hM;
    float *d_px;
    float2 *d_U;
    float *d_aux;
    bool profiling;
};

struct ModeCreate {
    int mode;
    int w, h, W, H, nP, M, nF;
    float radius, tSS, tSO;
    const esrrf_params *p;
};

struct ModeRgc {
    int frame0, nframes, count0, nFrameForSRRF;
    float sens;
    int sens_int, int_weighting;
    float radius, tSS, tSO;
    const float *drift;
    bool drift_zero;
    float *out, *prev;
};

int mode_create(const ModeCreate &c, ModeState &ms, float **d_px, float2 **d_U, float **d_aux,
                size_t *aux_bytes);
int mode_prepare(int mode, const ModeCtx &c, ModeState &ms, const void *src, int dtype, int n);
int mode_rgc(int mode, const ModeCtx &c, ModeState &ms, const ModeRgc &r);
int mode_read_gradients(int mode, const ModeCtx &c, ModeState &ms, float *out);
long long mode_active_samples(int mode, const ModeState &ms);
void free_mode_state(ModeState &ms);

}  // namespace esrrf
