// modes.h -- FHT (Fourier-upsampled) and 3-D (multi-focus) pipelines of the engine.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/esrrf.h"
#include "esrrf_internal.h"
#include "fht_internal.h"
#include "variants.h"

namespace esrrf {

struct ModeState {
    int mode = 0;                 // 1 = FHT, 2 = 3-D
    std::string err;
    double prof_up_ms = 0.0;      // upsampling kernel time (FHT)
    long prof_up_n = 0;
    long long active = 0;         // gated RGC samples per frame (geometry only)
    int nF = 0;
    int wM = 0, hM = 0;
    // FHT
    float *d_rawf = nullptr;      // converted raw frames of the load [nF][h][w]
    FhtPlan plan{};
    std::vector<Tap> h_taps;      // unshifted-window tap table (magnified grid)
    std::vector<int> h_tapk;
    Tap *d_taps = nullptr;
    int *d_tapk = nullptr;
    int ntaps = 0;
    float tap_wsum = 0.0f;
    int fht_R = 0, fht_Rhi = 0, fht_TW = 0, fht_pitch = 0;
    // 3-D
    float *d_st = nullptr;        // [3][nP] -theta (rad), shiftX, shiftY (LiveSRRF_CL.java:440-445)
    int *d_cell = nullptr;        // [2][nP] montage cell origins
    float *d_icf = nullptr;       // [nP] intensity scaling
    float bg = 0.0f;
    float *d_planes = nullptr;    // reshaped planes [nF][nP][h][w]
    float *d_g3 = nullptr;        // 2-point gradients [3][nF*nP][h][w]
    float4 *d_u3 = nullptr;       // unit tap gradients [nF*nP][2h][2w]
    float *d_lat = nullptr;       // lateral intensities [nF][nP][hM][wM]
    int n_loaded = 0;             // frames of the current load (gradient plane stride)
    int Mz = 1;                   // axial magnification
    // tiled 3-D RGC tables (frames without drift)
    float4 *d_tapv3 = nullptr;
    int *d_tapo3 = nullptr;
    int *d_tapk3 = nullptr;
    Phase3 *d_ph3 = nullptr;
    Rgc3Tables t3{};
    bool tiled3 = false;
    size_t lds_limit = 0;         // dynamic LDS one block may use on the device
    KernelVariants var;           // alternative kernels (read when the engine was created)
    const char *rgc_kernel = "";  // RGC kernel of the last launch
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
    int Mz;                       // 3-D: axial magnification (planes per input plane)
    float radius, tSS, tSO;
    const esrrf_params *p;
    KernelVariants var;
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
// work a mode does per launch before its RGC kernel (3-D: the lateral interpolation), so
// that the RGC can be timed alone
int mode_pre_rgc(int mode, const ModeCtx &c, ModeState &ms, const ModeRgc &r);
int mode_rgc(int mode, const ModeCtx &c, ModeState &ms, const ModeRgc &r);
int mode_read_gradients(int mode, const ModeCtx &c, ModeState &ms, float *out);
// 3-D mode: reshapeImageStack3D of n montage frames at src (device) into out (host)
int mode_reshape_3d(const ModeCtx &c, ModeState &ms, const void *src, int dtype, int n,
                    float *out);
long long mode_active_samples(int mode, const ModeState &ms);
void free_mode_state(ModeState &ms);

// Device tap table of the FHT-mode RGC for one fwhm (the table mode_create builds for its
// own fwhm), for callers that run one FHT state with several fwhm values (the sweep).
struct FhtTaps {
    Tap *d_taps = nullptr;
    int *d_tapk = nullptr;
    int R = 0, Rhi = 0, TW = 0, pitch = 0, ntaps = 0;
    float wsum = 0.0f;
};
int fht_taps_create(float radius, float tSS, float tSO, FhtTaps &t);
void fht_taps_free(FhtTaps &t);

}  // namespace esrrf
