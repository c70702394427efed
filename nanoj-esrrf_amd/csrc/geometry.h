// geometry.h -- host-side tap tables of the tiled BIC RGC kernel.
#pragma once
#include <vector>

#include "esrrf_internal.h"

namespace esrrf {

struct BicGeometry {
    int R2lo = 0, R2hi = 0;   // window offsets of liveSRRF.cl:279,283
    int G0 = 0;               // even margin >= -R2lo of the LDS tile
    int TW = 0, TH = 0;       // LDS tile extent on the 2x grid
    int PW = 0, PH = 0, PP = 0;  // parity-plane width, height, pitch (float2)
    int xs_log2 = 3;
    int ncls_x = 1, ncls_y = 1;
    int max_off = 0;
    int max_class_taps = 0;   // taps of the largest x/y class (all its phases)
    std::vector<Tap> taps;
    std::vector<Phase> phases;
    std::vector<int> tapk;
    std::vector<float> tapWd;   // per tap dW/dtSS = W * 4 d^2 / tSS^2 (0 for padding)
};

bool build_bic_geometry(int w, int h, int M, float radius, float tSS, float tSO, BicGeometry &g);

// Union tables of nf <= SWEEP_NFG fwhm values of one magnification: the tile layout of the
// widest window, the taps of the widest gate; tap t carries W[f] of fwhm f (0 outside its
// gate) and phase wsum[f] sums fwhm f's weights in the same order as build_bic_geometry.
struct BicGeometryMulti {
    int R2lo = 0, R2hi = 0, G0 = 0, TW = 0, TH = 0, PW = 0, PH = 0, PP = 0;
    int xs_log2 = 3, ncls_x = 1, ncls_y = 1;
    int max_off = 0;
    int max_phase_taps = 0;   // longest tap list of one phase
    int nf = 0;
    std::vector<TapM> taps;
    std::vector<PhaseM> phases;
    std::vector<int> tapk;
};
bool build_bic_geometry_multi(int w, int h, int M, const float *radius, const float *tSS,
                              const float *tSO, int nf, BicGeometryMulti &g);
long long count_active_samples_bic(int w, int h, int M, float radius, float tSO);

}  // namespace esrrf

namespace esrrf {

// Host copy of the tiled 3-D RGC kernel's tables (Rgc3Tables without device pointers).
struct Geometry3 {
    int R2lo = 0, R2hi = 0, G0 = 0;
    int TW = 0, TH = 0, PW = 0, P = 0, PZ = 0, NZ = 0;
    int xs_log2 = 3, ncls_x = 1, ncls_y = 1, nZM = 0;
    int max_group_taps = 0;   // taps of the largest group of 4 consecutive phases
    int NZ2 = 0;              // planes of the union of z-slices 2k and 2k+1 (largest k)
    bool pair_ok = false;     // every tap's in-plane offset + lane base stays in its plane
    std::vector<float4> tapv;
    std::vector<int> tapo, tapk;
    std::vector<Phase3> phases;
};

// Tables for frames without drift; false when the window or tile exceeds the kernel's
// limits (the per-voxel kernel is used then).
bool build_3d_geometry(int w, int h, int nP, int M, int Mz, float radius, float tSS,
                       float tSO, Geometry3 &g);

}  // namespace esrrf
