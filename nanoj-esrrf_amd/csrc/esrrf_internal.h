// esrrf_internal.h -- shared host/device declarations of the eSRRF engine (not public ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace esrrf {

// Zero-weight taps at the end of the BIC tap table: the tiled kernel's prefetch of the next
// chunks of a list may read up to this many entries past a class's last tap.
constexpr int TAP_SPARE = 8;

// One radial-gradient-convergence sample ("tap") of a magnified pixel.
//   off : float2 index of the tap's unit gradient inside the block's LDS tile,
//         relative to the lane's base index (see tile layout in k_bic.hip)
//   a,b : unit vector (dx, dy)/d from the pixel centre to the sample point
//   W   : distance weight (d * exp(-d^2 / tSS))^4  (liveSRRF.cl:296-297)
struct Tap {
    int off;
    float a;
    float b;
    float W;
};

// Taps of several fwhm values at once (the parameter sweep's emit pass): the union of their
// gates in the reference's j-then-i order, each tap with the distance weight of every fwhm
// of the group (0 where that fwhm's gate excludes it).
constexpr int SWEEP_NFG = 5;   // fwhm values per union table
struct TapM {
    int off;
    float a, b;
    float W[SWEEP_NFG];
};
static_assert(sizeof(TapM) == 32, "TapM is two float4");

struct PhaseM {
    int first_tap, ntaps;    // into the TapM array; ntaps a multiple of 4
    int du, dv;
    float wx[4], wy[4];
    int sx, sy, pad0, pad1;
    float wsum[8];           // per fwhm: sum of W over its gated taps, reference order
};

// Geometry of one (x-class, y-class, sub-pixel phase) of a magnified pixel.
struct Phase {
    int first_tap;  // index into the Tap array
    int ntaps;
    float wsum;     // sum of W over all taps, in the reference's j-then-i order
    int du, dv;     // bicubic base u0 - X, v0 - Y of the intensity sample
    float wx[4];    // Catmull-Rom weights along x for u0-1 .. u0+2
    float wy[4];
    int sx, sy;     // sub-pixel phase (xM mod M, yM mod M)
    float wsumd;    // sum of dW/dtSS over all taps (same order as wsum)
    int ntaps2;     // gated taps rounded up to an even count (<= ntaps): the taps beyond it
                    // are zero-weight padding, which adds W * h = -0 to a sum
};

// Launch parameters of the BIC 2-D radial-gradient-convergence kernel.
struct BicRgcArgs {
    const float2 *U;       // [nF][2h][2w] unit tap gradients
    const float *px;       // [nF][h][w] raw frames (float)
    const Tap *taps;
    const Phase *phases;   // [ncls_y][ncls_x][M*M]
    const int *tapk;       // per tap: (kx_rel + 128) | (ky_rel + 128) << 8
    float *out;            // [4][hM][wM]
    float *prev;           // [hM][wM]
    int w, h, M, wM, hM;
    int frame0, nframes;   // frames [frame0, frame0 + nframes) of the load
    int first_init;        // 1: the launch's first frame starts the SR frame
    float rn, rn1;         // 1/nFrameForSRRF, 1/(nFrameForSRRF - 1)
    float sens;
    int sens_int;          // 1..4 exact integer power, 0 = general powf
    int int_weighting;
    int tac_prev_scaled;   // 1: Prev stores t/(n-1) (BIC quirk)
    int G0, PP, PH, TW, TH;
    int R2lo, R2hi;
    int xs_log2, ncls_x, ncls_y;
    int tiles_x, tiles_y;
    int nw;                // waves per block
    int zero;              // always 0 (defeats loop-invariant hoisting of phase tables)
    int stage;             // 1: stage the next frame's tiles in registers (k_bic_rgc)
    int pair_direct;       // 1: the frame-pair kernel fills the next pair's tiles after the
                           //    current pair instead of staging them in registers
    int taps_in_lds;       // 1: stage each block's tap tables in LDS
    int max_class_taps;    // largest tap-table size of one x/y class (LDS sizing)
    int frames_per_pass;   // 4: four-frame kernel (mode 0), 2: frame-pair kernel, else 1
    // kernel variant: 0 = accumulate the four cumulants (liveSRRF.cl:363-393);
    // 1 = emit the normalised CGLH (before the sensitivity power) and the intensity of
    //     every pixel and frame, phase-major [frame][M*M][h][w], into emitC / emitV;
    // 2 = accumulate AVG, dAVG/dtSS, dAVG/dS and INT (out channels 0..3) for the RSE adjoint
    int mode;
    float *emitC, *emitV;
    const float *tapWd;    // per tap dW/dtSS (mode 2)
};

struct GenericRgcArgs {
    const float2 *U;
    const float *px;
    const float *drift;    // [2][nFrameForSRRF]
    float *out;
    float *prev;
    int w, h, M, wM, hM;
    int frame0, nframes, count0;
    int nFrameForSRRF;
    float rn, rn1;
    float sens;
    int sens_int;
    int int_weighting;
    int tac_prev_scaled;
    float radius, tSS, tSO;
};

// Geometry of one (y-class, x-class, magnified z-slice, lateral sub-pixel phase) of the
// tiled 3-D RGC kernel.  Its taps are listed in live3DSRRF.cl's j, i, k loop order.
struct Phase3 {
    int first_tap;  // index into tapv / tapo / tapk
    int ntaps;      // a multiple of 4 (zero-weight padding at the end)
    float wsum;     // sum of W over the phase's taps, in the reference's order
    int zlo, nz;    // planes [zlo, zlo + nz) of the z-slice's LDS tile
    int sx, sy;     // lateral sub-pixel phase (xM mod M, yM mod M)
    int pad;
};

// Tap tables and LDS tile layout of the tiled 3-D RGC kernel (built once per engine).
//   tapv : (ex, ey, ez, W) unit direction from the voxel centre to the sample point and
//          its distance weight (d exp(-d^2 / tSS))^4 (live3DSRRF.cl:383-384)
//   tapo : float4 index of the sample's unit gradient in the LDS tile, relative to the
//          lane's base index 2 ly P + lx
//   tapk : (kx_rel + 128) | (ky_rel + 128) << 8, the sample's 2x-grid offset from (2X, 2Y)
struct Rgc3Tables {
    const float4 *tapv;
    const int *tapo;
    const int *tapk;
    const Phase3 *phases;   // [ncls_y][ncls_x][nP M][M M]
    int R2lo, R2hi, G0;
    int TW, TH;             // tile extent on the 2x grid (columns, rows)
    int PW, P, PZ, NZ;      // half-row width, row pitch, plane pitch (float4), planes
    int xs_log2, ncls_x, ncls_y, nZM;
    int gtaps;              // LDS tap slots per block (largest group of 4 phases + spares)
    int NZ2;                // planes of a two-slice block's tile (union of the slices' planes)
};

// Host helpers implemented in the .hip files (all launches on `s`).
hipError_t launch_convert_frames(const void *src, int dtype, float *dst, long n, hipStream_t s);
// kind 4: k_bic_prep4 (four frames per block, the default); 2: k_bic_prep2; 1: k_bic_prep
hipError_t launch_bic_prep(const float *px, float2 *U, int w, int h, int nF, hipStream_t s,
                           int kind = 4);
hipError_t launch_bic_gint(const float *px, float *gxi, float *gyi, int w, int h, int f,
                           hipStream_t s);
// *kname (when given) receives the name of the kernel launched.
hipError_t launch_bic_rgc(const BicRgcArgs &a, int ppw, hipStream_t s,
                          const char **kname = nullptr);
size_t bic_pair_lds_bytes(const BicRgcArgs &a, bool tsm);
// four frames per pass (k_bic_rgc4): LDS bytes and whether the geometry fits the kernel
size_t bic_quad_lds_bytes(const BicRgcArgs &a);
bool bic_quad_fits(const BicRgcArgs &a);
// The sweep's emit pass over a union table of nfg <= SWEEP_NFG fwhm values (k_bic_emit_multi):
// a.nw waves x ppw (1 or 2) phases per block, pg_taps = the largest tap count of one block's
// phases (LDS sizing), CGLH of fwhm f at a.emitC + f * cstride.
hipError_t launch_bic_emit_multi(const BicRgcArgs &a, int ppw, const TapM *taps,
                                 const PhaseM *phases, const int *tapk, int nfg, int pg_taps,
                                 long cstride, hipStream_t s);
size_t bic_emit_multi_lds_bytes(const BicRgcArgs &a, int pg_taps);   // LDS of the frame-pair kernel
hipError_t launch_generic_rgc(const GenericRgcArgs &a, hipStream_t s);
// nbatch images `ostride` floats apart (MP maps `mstride` floats apart) in one launch.
// MP maps: nP input planes magnified M laterally and Mz axially (2-D: nP = Mz = 1).
hipError_t launch_var(float *out, long nvox, hipStream_t s, int nbatch = 1, long ostride = 0);
hipError_t launch_mp_map(const float *out, float *mp, int w, int h, int nP, int M, int Mz,
                         hipStream_t s, int nbatch = 1, long ostride = 0, long mstride = 0);
hipError_t launch_mp_correct(float *out, const float *mp, int w, int h, int nP, int M, int Mz,
                             hipStream_t s, int nbatch = 1, long ostride = 0, long mstride = 0);
hipError_t launch_nan_to_zero_copy(const float *src, float *dst, long n, hipStream_t s);

}  // namespace esrrf
