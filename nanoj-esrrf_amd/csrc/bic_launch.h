// bic_launch.h -- device tap tables, launch arguments, the default launch shape and the
// autotuner of the tiled BIC RGC kernel; shared by the engine, the parameter sweep and
// the RSE adjoint.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "esrrf_internal.h"
#include "geometry.h"

namespace esrrf {

// Launch shape of k_bic_rgc: nw waves per block, ppw sub-pixel phases per wave,
// tap tables staged in LDS or read from the global tables.
struct BicShape {
    int nw = 1, ppw = 1;
    int taps_in_lds = 0;
    int fp = 1;           // frames per pass: 4 = four-frame kernel (mode 0, taps in LDS,
                          // ppw <= 4, tile fits), 2 = frame-pair kernel (ppw <= 4)
    float ms = 0.0f;      // measured time of the autotuner's sample launch (0: not tuned)
};

// Geometry of one (w, h, M, fwhm) plus its device copy.
struct BicTables {
    BicGeometry geo;
    Tap *d_taps = nullptr;
    Phase *d_phases = nullptr;
    int *d_tapk = nullptr;
    float *d_tapWd = nullptr;
    float radius = 0, tSS = 0, tSO = 0;
    bool ok = false;      // geometry fits the tiled kernel
};

// Build the tables for frames of w x h, magnification M and the constants of fwhm
// (LiveSRRF_CL.java:304-329) and upload them (synchronous copies: not for capture).
hipError_t bic_tables_create(int w, int h, int M, float fwhm, BicTables &t);
void bic_tables_free(BicTables &t);

// Default shape: waves per block a multiple of 4 (a block's waves are dealt round-robin
// to the CU's four SIMDs), about 4 phases per wave, at most 16 waves and 8 phases per wave;
// tap tables in LDS when a class's tables stay within 48 KiB (adj: with dW/dtSS); frame
// pairs wherever that kernel applies.
BicShape bic_default_shape(const BicTables &t, int M, bool adj = false);
// All shapes the kernel supports for M (the autotuner's candidates), frame-pair shapes
// included where their LDS tiles and tables fit in 64 KiB (not for the adjoint).
std::vector<BicShape> bic_candidate_shapes(const BicTables &t, int M, bool adj = false);
// Whether the kernels run shape s (any wave count, not only the candidates) for this
// geometry.
bool bic_shape_supported(const BicTables &t, int M, const BicShape &s, bool adj = false);

// Arguments common to every launch over unshifted frames (out, prev, frames, cumulant
// constants and mode are filled by the caller).
BicRgcArgs bic_args(const BicTables &t, const BicShape &sh, const float2 *U, const float *px,
                    int w, int h, int M);

// Time every candidate shape on `a` (reps launches each after one warm-up) and return the
// fastest; the launches write whatever `a` points at, so callers pass scratch outputs.
hipError_t bic_autotune(const BicTables &t, BicRgcArgs a, int M, int reps, hipStream_t s,
                        BicShape *best, std::vector<BicShape> *all);

int sens_exact_int(float s);

// Union tables of up to SWEEP_NFG fwhm values of one magnification (the sweep's emit pass,
// k_bic_emit_multi) and their device copy.
struct BicTablesMulti {
    BicGeometryMulti geo;
    TapM *d_taps = nullptr;
    PhaseM *d_phases = nullptr;
    int *d_tapk = nullptr;
    int nfg = 0;
    bool ok = false;      // geometry fits the kernel
};
hipError_t bic_multi_tables_create(int w, int h, int M, const float *fwhm, int nfg,
                                   BicTablesMulti &t);
void bic_multi_tables_free(BicTablesMulti &t);

// Launch shape of the emit pass: nw waves x ppw phases per block; pg_taps the longest tap
// list of one block's phase group (LDS sizing).
struct EmitShape {
    int nw = 8, ppw = 1, pg_taps = 0;
    float ms = 0.0f;
};
// Shapes whose LDS stays within 64 KiB (two blocks per CU), default first.
std::vector<EmitShape> emit_candidate_shapes(const BicTablesMulti &t, int M);
BicRgcArgs bic_args_multi(const BicTablesMulti &t, const EmitShape &sh, const float2 *U,
                          const float *px, int w, int h, int M);
// Time every candidate on `a` (reps launches after a warm-up) and keep the fastest.
hipError_t emit_autotune(const BicTablesMulti &t, BicRgcArgs a, int M, long cstride, int reps,
                         hipStream_t s, EmitShape *best);

}  // namespace esrrf
