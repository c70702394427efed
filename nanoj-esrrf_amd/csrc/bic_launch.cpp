// bic_launch.cpp -- device tap tables, default launch shape and autotuner of the tiled BIC
// RGC kernel.
#include "bic_launch.h"

#include <algorithm>

#include "../../include/esrrf.h"

namespace esrrf {

int sens_exact_int(float s)
{
    for (int k = 1; k <= 4; k++)
        if (s == (float)k) return k;
    return 0;
}

void bic_tables_free(BicTables &t)
{
    void *ptrs[] = {t.d_taps, t.d_phases, t.d_tapk, t.d_tapWd};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    t.d_taps = nullptr;
    t.d_phases = nullptr;
    t.d_tapk = nullptr;
    t.d_tapWd = nullptr;
    t.ok = false;
}

hipError_t bic_tables_create(int w, int h, int M, float fwhm, BicTables &t)
{
    float sigma = 0.0f;
    esrrf_bic_constants(fwhm, &sigma, &t.radius, &t.tSS, &t.tSO);
    t.ok = M <= 8 && build_bic_geometry(w, h, M, t.radius, t.tSS, t.tSO, t.geo);
    if (!t.ok) return hipSuccess;
    const size_t nt = t.geo.taps.size(), np = t.geo.phases.size();
    hipError_t e = hipMalloc((void **)&t.d_taps, sizeof(Tap) * nt);
    if (e == hipSuccess) e = hipMalloc((void **)&t.d_phases, sizeof(Phase) * np);
    if (e == hipSuccess) e = hipMalloc((void **)&t.d_tapk, sizeof(int) * nt);
    if (e == hipSuccess) e = hipMalloc((void **)&t.d_tapWd, sizeof(float) * nt);
    if (e == hipSuccess)
        e = hipMemcpy(t.d_taps, t.geo.taps.data(), sizeof(Tap) * nt, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(t.d_phases, t.geo.phases.data(), sizeof(Phase) * np, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(t.d_tapk, t.geo.tapk.data(), sizeof(int) * nt, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(t.d_tapWd, t.geo.tapWd.data(), sizeof(float) * nt, hipMemcpyHostToDevice);
    if (e != hipSuccess) bic_tables_free(t);
    return e;
}

// tap tables (and dW/dtSS for the adjoint) of one class in at most 48 KiB of LDS, which
// leaves room for two blocks per CU
static bool tlds_fits(const BicTables &t, bool adj)
{
    return (size_t)(t.geo.max_class_taps + TAP_SPARE) *
               (sizeof(Tap) + (adj ? sizeof(float) : 0)) <=
           48 * 1024;
}

// LDS of the frame-pair kernel for this geometry within the 64 KiB a block may take by
// default
static bool pair_fits(const BicTables &t, bool tsm = false)
{
    BicRgcArgs a{};
    a.PP = t.geo.PP;
    a.PH = t.geo.PH;
    a.TW = t.geo.TW;
    a.max_class_taps = t.geo.max_class_taps;
    return bic_pair_lds_bytes(a, tsm) <= 64 * 1024;
}

// geometry of the four-frame kernel (parity-plane pitch, tile rows, LDS)
static bool quad_fits(const BicTables &t, int M)
{
    BicRgcArgs a{};
    a.PP = t.geo.PP;
    a.PH = t.geo.PH;
    a.TW = t.geo.TW;
    a.M = M;
    a.max_class_taps = t.geo.max_class_taps;
    return bic_quad_fits(a);
}

BicShape bic_default_shape(const BicTables &t, int M, bool adj)
{
    const int MM = M * M;
    int nw = MM <= 4 ? MM : std::min(16, 4 * (((MM + 3) / 4 + 3) / 4));
    if ((MM + nw - 1) / nw > 8) nw = std::min(16, (MM + 7) / 8);
    BicShape s;
    s.nw = nw;
    s.ppw = (MM + nw - 1) / nw;
    s.taps_in_lds = tlds_fits(t, adj) ? 1 : 0;
    // four frames per pass where the geometry fits that kernel (one tap read and one address
    // add per four samples), else frame pairs (per two samples)
    if (s.taps_in_lds && !adj && s.ppw <= 4 && quad_fits(t, M))
        s.fp = 4;
    else if (s.taps_in_lds && !adj && s.ppw <= 4 && pair_fits(t))
        s.fp = 2;
    return s;
}

bool bic_shape_supported(const BicTables &t, int M, const BicShape &s, bool adj)
{
    const int MM = M * M;
    if (s.nw < 1 || s.nw > 16 || s.ppw != (MM + s.nw - 1) / s.nw || s.ppw > 8) return false;
    if (s.taps_in_lds && !tlds_fits(t, adj)) return false;
    if (s.fp == 2) return !adj && s.ppw <= 4 && pair_fits(t, s.taps_in_lds == 0);
    if (s.fp == 4) return !adj && s.ppw <= 4 && s.taps_in_lds && quad_fits(t, M);
    return s.fp == 1;
}

std::vector<BicShape> bic_candidate_shapes(const BicTables &t, int M, bool adj)
{
    const int MM = M * M;
    std::vector<BicShape> out;
    std::vector<int> nws;
    if (MM <= 4) nws.push_back(MM);
    for (int nw = 4; nw <= 16; nw += 4)
        if (nw <= 4 * MM) nws.push_back(nw);
    for (int nw : nws) {
        const int ppw = (MM + nw - 1) / nw;
        if (ppw > 8 || ppw < 1) continue;
        for (int tl = tlds_fits(t, adj) ? 1 : 0; tl >= 0; tl--) {
            BicShape s;
            s.nw = nw;
            s.ppw = ppw;
            s.taps_in_lds = tl;
            out.push_back(s);
            // frame pairs, taps in LDS (tl 1) or through scalar loads (tl 0)
            if (!adj && ppw <= 4 && pair_fits(t, tl == 0)) {
                s.fp = 2;
                out.push_back(s);
            }
            // four frames per pass (taps in LDS)
            if (!adj && ppw <= 4 && tl == 1 && quad_fits(t, M)) {
                s.fp = 4;
                out.push_back(s);
            }
        }
    }
    return out;
}

BicRgcArgs bic_args(const BicTables &t, const BicShape &sh, const float2 *U, const float *px,
                    int w, int h, int M)
{
    BicRgcArgs a{};
    a.U = U;
    a.px = px;
    a.taps = t.d_taps;
    a.phases = t.d_phases;
    a.tapk = t.d_tapk;
    a.tapWd = t.d_tapWd;
    a.w = w;
    a.h = h;
    a.M = M;
    a.wM = w * M;
    a.hM = h * M;
    a.G0 = t.geo.G0;
    a.PP = t.geo.PP;
    a.PH = t.geo.PH;
    a.TW = t.geo.TW;
    a.TH = t.geo.TH;
    a.R2lo = t.geo.R2lo;
    a.R2hi = t.geo.R2hi;
    a.xs_log2 = t.geo.xs_log2;
    a.ncls_x = t.geo.ncls_x;
    a.ncls_y = t.geo.ncls_y;
    a.tiles_x = (w + 7) / 8;
    a.tiles_y = (h + 7) / 8;
    a.nw = sh.nw;
    a.stage = 0;
    a.taps_in_lds = sh.taps_in_lds;
    a.frames_per_pass = sh.fp;
    a.max_class_taps = t.geo.max_class_taps;
    a.tac_prev_scaled = 1;
    return a;
}

hipError_t bic_autotune(const BicTables &t, BicRgcArgs a, int M, int reps, hipStream_t s,
                        BicShape *best, std::vector<BicShape> *all)
{
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate(&e0);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        return e;
    }
    *best = bic_default_shape(t, M, a.mode == 2);
    best->ms = 0.0f;
    bool have = false;
    for (BicShape c : bic_candidate_shapes(t, M, a.mode == 2)) {
        a.nw = c.nw;
        a.taps_in_lds = c.taps_in_lds;
        a.frames_per_pass = c.fp;
        e = launch_bic_rgc(a, c.ppw, s);   // warm-up
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_bic_rgc(a, c.ppw, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) break;
        c.ms = ms / (float)std::max(reps, 1);
        if (all) all->push_back(c);
        if (!have || c.ms < best->ms) {
            *best = c;
            have = true;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return e;
}

void bic_multi_tables_free(BicTablesMulti &t)
{
    void *ptrs[] = {t.d_taps, t.d_phases, t.d_tapk};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    t.d_taps = nullptr;
    t.d_phases = nullptr;
    t.d_tapk = nullptr;
    t.ok = false;
}

hipError_t bic_multi_tables_create(int w, int h, int M, const float *fwhm, int nfg,
                                   BicTablesMulti &t)
{
    float radius[SWEEP_NFG], tss[SWEEP_NFG], tso[SWEEP_NFG];
    for (int f = 0; f < nfg; f++) {
        float sigma = 0.0f;
        esrrf_bic_constants(fwhm[f], &sigma, &radius[f], &tss[f], &tso[f]);
    }
    t.nfg = nfg;
    t.ok = M <= 8 && nfg >= 1 && nfg <= SWEEP_NFG &&
           build_bic_geometry_multi(w, h, M, radius, tss, tso, nfg, t.geo);
    if (!t.ok) return hipSuccess;
    const size_t nt = t.geo.taps.size(), np = t.geo.phases.size();
    hipError_t e = hipMalloc((void **)&t.d_taps, sizeof(TapM) * nt);
    if (e == hipSuccess) e = hipMalloc((void **)&t.d_phases, sizeof(PhaseM) * np);
    if (e == hipSuccess) e = hipMalloc((void **)&t.d_tapk, sizeof(int) * nt);
    if (e == hipSuccess)
        e = hipMemcpy(t.d_taps, t.geo.taps.data(), sizeof(TapM) * nt, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(t.d_phases, t.geo.phases.data(), sizeof(PhaseM) * np,
                      hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(t.d_tapk, t.geo.tapk.data(), sizeof(int) * nt, hipMemcpyHostToDevice);
    if (e != hipSuccess) bic_multi_tables_free(t);
    return e;
}

// longest tap run of one block's phases [g PG, (g+1) PG) over every x/y class
static int group_taps(const BicTablesMulti &t, int M, int pg)
{
    const int MM = M * M;
    int best = 0;
    const int ncls = t.geo.ncls_x * t.geo.ncls_y;
    for (int c = 0; c < ncls; c++) {
        const PhaseM *P = t.geo.phases.data() + (size_t)c * MM;
        for (int p0 = 0; p0 < MM; p0 += pg) {
            const int pl = std::min(p0 + pg, MM) - 1;
            best = std::max(best, P[pl].first_tap + P[pl].ntaps - P[p0].first_tap);
        }
    }
    return best;
}

BicRgcArgs bic_args_multi(const BicTablesMulti &t, const EmitShape &sh, const float2 *U,
                          const float *px, int w, int h, int M)
{
    BicRgcArgs a{};
    a.U = U;
    a.px = px;
    a.w = w;
    a.h = h;
    a.M = M;
    a.wM = w * M;
    a.hM = h * M;
    a.G0 = t.geo.G0;
    a.PP = t.geo.PP;
    a.PH = t.geo.PH;
    a.TW = t.geo.TW;
    a.TH = t.geo.TH;
    a.R2lo = t.geo.R2lo;
    a.R2hi = t.geo.R2hi;
    a.xs_log2 = t.geo.xs_log2;
    a.ncls_x = t.geo.ncls_x;
    a.ncls_y = t.geo.ncls_y;
    a.tiles_x = (w + 7) / 8;
    a.tiles_y = (h + 7) / 8;
    a.nw = sh.nw;
    a.mode = 1;
    return a;
}

std::vector<EmitShape> emit_candidate_shapes(const BicTablesMulti &t, int M)
{
    std::vector<EmitShape> out;
    const int MM = M * M;
    for (int ppw = 1; ppw <= 2; ppw++) {
        for (int nw = 4; nw <= 16; nw += 4) {
            if ((nw - 4) * ppw >= MM) break;   // smaller blocks already hold every phase
            EmitShape s;
            s.nw = nw;
            s.ppw = ppw;
            s.pg_taps = group_taps(t, M, nw * ppw);
            BicRgcArgs a = bic_args_multi(t, s, nullptr, nullptr, 8, 8, M);
            if (bic_emit_multi_lds_bytes(a, s.pg_taps) <= 64 * 1024) out.push_back(s);
        }
    }
    if (out.empty()) {   // one phase per wave and block: the smallest footprint
        EmitShape s;
        s.nw = 1;
        s.ppw = 1;
        s.pg_taps = group_taps(t, M, 1);
        out.push_back(s);
    }
    return out;
}

hipError_t emit_autotune(const BicTablesMulti &t, BicRgcArgs a, int M, long cstride, int reps,
                         hipStream_t s, EmitShape *best)
{
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate(&e0);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        return e;
    }
    const std::vector<EmitShape> cands = emit_candidate_shapes(t, M);
    *best = cands[0];
    bool have = false;
    for (EmitShape c : cands) {
        a.nw = c.nw;
        e = launch_bic_emit_multi(a, c.ppw, t.d_taps, t.d_phases, t.d_tapk, t.nfg, c.pg_taps,
                                  cstride, s);   // warm-up
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int r = 0; r < reps && e == hipSuccess; r++)
            e = launch_bic_emit_multi(a, c.ppw, t.d_taps, t.d_phases, t.d_tapk, t.nfg,
                                      c.pg_taps, cstride, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) break;
        c.ms = ms / (float)std::max(reps, 1);
        if (!have || c.ms < best->ms) {
            *best = c;
            have = true;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return e;
}

}  // namespace esrrf
