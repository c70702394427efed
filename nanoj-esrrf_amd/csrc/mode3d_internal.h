// mode3d_internal.h -- the 3-D (multi-focus) pipeline: host entry points and kernel
// launch parameters.
#pragma once
#include "modes.h"

namespace esrrf {

struct Prep3dArgs {
    int W, H;              // montage frame
    int w, h, nP;          // one plane
    const int *cell;       // [2][nP] montage cell origin x, y (LiveSRRF_CL.java:973-979)
    const float *ic;       // [nP] intensity scaling
    float bg;              // background level
    const float *st;       // [3][nP] -theta (rad), shiftX, shiftY (LiveSRRF_CL.java:440-445)
    float *planes;         // [nF][nP][h][w] reshaped
    float *aligned;        // [nF][nP][h][w] aligned
    float *g;              // [3][nF*nP][h][w] 2-point gradients
    float4 *U;             // [nF*nP][2h][2w] unit tap gradients
};

struct Rgc3dArgs {
    const float4 *U;
    const float *px;       // aligned planes
    float *L;              // [nframes][nP][hM][wM] lateral intensities (scratch)
    const float *drift;    // [2][nFor]
    float *out, *prev;
    int w, h, nP, M, wM, hM;
    int Mz;                // axial magnification: nP * Mz output planes, zc = (zM + 0.5) / Mz
    int frame0, nframes, count0, nFor;
    float sens;
    int sens_int, int_weighting;
    float radius, tSS, tSO;
};

hipError_t launch3d_prepare(const void *src, int dtype, const Prep3dArgs &p, int n,
                            hipStream_t s);
// k3d_reshape alone: montage cells of n frames -> p.planes
hipError_t launch3d_reshape(const void *src, int dtype, const Prep3dArgs &p, int n,
                            hipStream_t s);
hipError_t launch3d_gint(const float *g, float *out, int w, int h, int nP, long gn,
                         hipStream_t s);
// The tiled kernel (k3d_rgc3) runs for frames without drift when `t` is given and its LDS
// fits `lds_limit`; *kname receives the name of the kernel launched.
// kern: 3 tiled (default), 2 / 1 the per-voxel kernels; split: waves per phase in the tiled
// kernel (2 default)
// lateral bicubic of the frames' planes at the magnified grid (k3d_lateral), before the RGC
hipError_t launch3d_lateral(const Rgc3dArgs &a, hipStream_t s);
hipError_t launch3d_rgc(const Rgc3dArgs &a, bool drift_zero, const Rgc3Tables *t,
                        size_t lds_limit, hipStream_t s, const char **kname, int kern = 3,
                        int split = 2, int zpair = 1);
size_t rgc3_lds_bytes(const Rgc3Tables &t);
size_t rgc3z_lds_bytes(const Rgc3Tables &t);   // two-slice tiled kernel
hipError_t launch3d_count(int w, int h, int nP, int M, int Mz, float radius, float tSO,
                          unsigned long long *d_total, hipStream_t s);

int mode3d_create(const ModeCreate &c, ModeState &ms, float **d_px, float2 **d_U);
int mode3d_prepare(const ModeCtx &c, ModeState &ms, const void *src, int dtype, int n);
int mode3d_lateral(const ModeCtx &c, ModeState &ms, const ModeRgc &r);
int mode3d_rgc(const ModeCtx &c, ModeState &ms, const ModeRgc &r);
int mode3d_read_gradients(const ModeCtx &c, ModeState &ms, float *out);
// reshapeImageStack3D: n montage frames at src (device) -> out (host) [n][nP][h][w]
int mode3d_reshape(const ModeCtx &c, ModeState &ms, const void *src, int dtype, int n,
                   float *out);

}  // namespace esrrf
