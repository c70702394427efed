// sweep.cpp -- the parameter sweep of gui/ParametersSweep_.java:328-519 as one fused device
// sequence.
//
// The reference re-initialises LiveSRRF_CL (a program rebuild) for every
// (frame count, sensitivity, fwhm) and pushes the first nf frames one at a time
// (:452-458, :800-844).  Here:
//   * the frames go to the device once and their unit tap gradients are computed once
//     (they depend on neither magnification, fwhm nor sensitivity);
//   * per magnification and group of up to five fwhm values one RGC launch over the largest
//     frame count emits the normalised CGLH of every fwhm of the group and the intensity of
//     every pixel and frame (one angular term per sample serves every fwhm: the fwhm only
//     sets the weights and the gate, k_bic_emit_multi);
//   * one accumulation launch turns them into the cumulants of every (frame count,
//     sensitivity) configuration;
//   * VAR and MP correction finish each configuration (liveSRRF.cl:427-505).
// The whole device sequence can be captured once in a hipGraph and replayed, and the RGC
// launch shape of each (magnification, fwhm) can be picked by the autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "../../include/esrrf.h"
#include "bic_launch.h"
#include "esrrf_internal.h"
#include "fht_internal.h"
#include "modes.h"
#include "sweep_internal.h"
#include "variants.h"

using namespace esrrf;

struct esrrf_sweep {
    std::mutex mu;
    std::string err;
    int dev = 0;
    hipStream_t own_stream = nullptr, stream = nullptr;
    int w = 0, h = 0;
    std::vector<int> Ms;
    std::vector<float> fwhm, sens;
    std::vector<int> nfs;
    int nfmax = 0;
    int int_weighting = 1, do_mp = 1, autotune = 0, use_graph = 0;
    // device data
    void *d_in = nullptr;          // frames as handed over (raw dtype), nfmax frames
    float *d_px = nullptr;         // [nfmax][h][w]
    float2 *d_U = nullptr;         // [nfmax][2h][2w]
    float *d_C = nullptr, *d_V = nullptr;   // emitted CGLH [fwhm of group][..] / intensity
    long cstride = 0;              // floats between the CGLH planes of two fwhm
    bool multi = true;             // one emit launch per group of fwhm (else per fwhm)
    // FHT mode: per magnification the upsampled frames and their gradients, per fwhm the
    // magnified-space tap table
    int fht = 0;
    std::vector<ModeState> fms;                 // [iM]
    std::vector<float *> f_px, f_aux;           // [iM]
    std::vector<float2 *> f_U;                  // [iM]
    std::vector<std::vector<FhtTaps>> ftaps;    // [iM][iF]
    float *d_dummy = nullptr;      // out/prev of the emit launches (never written)
    std::vector<std::vector<BicTables>> tabs;   // [iM][iF]
    std::vector<std::vector<BicShape>> shape;   // [iM][iFwhm]
    std::vector<std::vector<BicTablesMulti>> mtabs;   // [iM][group of SWEEP_NFG fwhm]
    std::vector<std::vector<EmitShape>> mshape;       // [iM][group]
    std::vector<float *> d_out;    // [iM] configurations [nf][S][fwhm][4][hM][wM]
    std::vector<float *> d_mp;     // [iM] MP maps [nf][S][fwhm][3][M][M]
    float *d_tmp = nullptr;        // read-back staging
    size_t tmp_floats = 0;
    // graph of the device sequence (valid for one dtype)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_dtype = -1;
    bool tuned = false;
    bool have_result = false;
};

namespace {

int sfail(esrrf_sweep *s, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (s) s->err = buf;
    return code;
}

#define SCHK(s, x)                                                                         \
    do {                                                                                   \
        hipError_t _st = (x);                                                              \
        if (_st != hipSuccess)                                                             \
            return sfail((s), ESRRF_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(_st)); \
    } while (0)

void sweep_free(esrrf_sweep *s)
{
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    void *ptrs[] = {s->d_in, s->d_px, s->d_U, s->d_C, s->d_V, s->d_dummy, s->d_tmp};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &v : s->tabs)
        for (auto &t : v) bic_tables_free(t);
    for (auto &v : s->mtabs)
        for (auto &t : v) bic_multi_tables_free(t);
    for (auto &v : s->ftaps)
        for (auto &t : v) fht_taps_free(t);
    for (auto &m : s->fms) free_mode_state(m);
    for (float *p : s->f_px)
        if (p) (void)hipFree(p);
    for (float2 *p : s->f_U)
        if (p) (void)hipFree(p);
    for (float *p : s->d_out)
        if (p) (void)hipFree(p);
    for (float *p : s->d_mp)
        if (p) (void)hipFree(p);
    if (s->own_stream) (void)hipStreamDestroy(s->own_stream);
}

size_t n_conf(const esrrf_sweep *s) { return s->nfs.size() * s->sens.size() * s->fwhm.size(); }

// The device sequence of one run: frames in d_in -> every configuration's result.
// FHT mode: per magnification the frames are upsampled and their gradients formed once
// (mode_prepare), then per fwhm one RGC launch emits the normalised CGLH of every magnified
// pixel and frame and one accumulation launch forms every (frame count, sensitivity)
// configuration with the FHT cumulants (liveSRRF_noInterpolation.cl:358-390); VAR
// finishes them (no MP correction in FHT mode, LiveSRRF_CL.java:762).
int enqueue_run_fht(esrrf_sweep *s, int dtype)
{
    const hipStream_t st = s->stream;
    const long wh = (long)s->w * s->h;
    const int nS = (int)s->sens.size(), nF = (int)s->fwhm.size(), nQ = (int)s->nfs.size();
    for (size_t im = 0; im < s->Ms.size(); im++) {
        const int M = s->Ms[im];
        const long whM = wh * M * M;
        ModeCtx c = {st, s->w, s->h, s->w, s->h, 1, M, s->w * M, s->h * M,
                     s->f_px[im], s->f_U[im], s->f_aux[im], false};
        int rc = mode_prepare(1, c, s->fms[im], s->d_in, dtype, s->nfmax);
        if (rc != ESRRF_OK)
            return sfail(s, rc, "FHT upsampling: %s", s->fms[im].err.c_str());
        for (int fi = 0; fi < nF; fi++) {
            const FhtTaps &t = s->ftaps[im][fi];
            FhtRgcArgs a{};
            a.U = s->f_U[im];
            a.px = s->f_px[im];
            a.taps = t.d_taps;
            a.tapk = t.d_tapk;
            a.W = s->w * M;
            a.H = s->h * M;
            a.frame0 = 0;
            a.nframes = s->nfmax;
            a.first_init = 1;
            a.rn = a.rn1 = a.sens = 1.0f;
            a.sens_int = 1;
            a.int_weighting = s->int_weighting;
            a.R = t.R;
            a.Rhi = t.Rhi;
            a.pair = s->fms[im].var.fht_pair;
            a.TW = t.TW;
            a.TH = t.TW;
            a.pitch = t.pitch;
            a.ntaps = t.ntaps;
            a.wsum = t.wsum;
            a.tiles_x = (a.W + 15) / 16;
            a.tiles_y = (a.H + 15) / 16;
            a.emitC = s->d_C;
            SCHK(s, launch_fht_rgc(a, st));
            SweepAccArgs q{};
            q.C = s->d_C;
            q.V = s->f_px[im];
            q.out = s->d_out[im];
            q.w = s->w;
            q.h = s->h;
            q.M = M;
            q.nframes = s->nfmax;
            q.nS = nS;
            q.n_fwhm = nF;
            q.i_fwhm = fi;
            q.nq = nQ;
            q.int_weighting = s->int_weighting;
            q.fht = 1;
            for (int k = 0; k < nQ; k++) {
                q.nf[k] = s->nfs[k];
                q.rn[k] = 1.0f / (float)s->nfs[k];
                q.rn1[k] = 1.0f / (float)(s->nfs[k] - 1);
            }
            for (int k = 0; k < nS; k++) {
                q.sens[k] = s->sens[k];
                q.sens_int[k] = sens_exact_int(s->sens[k]);
            }
            SCHK(s, launch_sweep_acc(q, st));
            float *out = s->d_out[im] + (long)fi * 4 * whM;
            SCHK(s, launch_var(out, whM, st, nQ * nS, (long)nF * 4 * whM));
        }
    }
    return ESRRF_OK;
}

int enqueue_run(esrrf_sweep *s, int dtype)
{
    if (s->fht) return enqueue_run_fht(s, dtype);
    const hipStream_t st = s->stream;
    const long wh = (long)s->w * s->h;
    SCHK(s, launch_convert_frames(s->d_in, dtype, s->d_px, wh * s->nfmax, st));
    SCHK(s, launch_bic_prep(s->d_px, s->d_U, s->w, s->h, s->nfmax, st));
    const int nS = (int)s->sens.size(), nF = (int)s->fwhm.size(), nQ = (int)s->nfs.size();
    for (size_t im = 0; im < s->Ms.size(); im++) {
        const int M = s->Ms[im];
        const long whM = wh * M * M;
        for (int fi = 0; fi < nF; fi++) {
            const int slot = s->multi ? fi % SWEEP_NFG : 0;
            if (s->multi && slot == 0) {
                // CGLH of every fwhm of the group, intensity once
                const int g = fi / SWEEP_NFG;
                const BicTablesMulti &t = s->mtabs[im][g];
                const EmitShape &sh = s->mshape[im][g];
                BicRgcArgs a = bic_args_multi(t, sh, s->d_U, s->d_px, s->w, s->h, M);
                a.emitC = s->d_C;
                a.emitV = s->d_V;
                a.frame0 = 0;
                a.nframes = s->nfmax;
                a.int_weighting = s->int_weighting;
                SCHK(s, launch_bic_emit_multi(a, sh.ppw, t.d_taps, t.d_phases, t.d_tapk, t.nfg,
                                              sh.pg_taps, s->cstride, st));
            } else if (!s->multi) {
                const BicTables &t = s->tabs[im][fi];
                BicRgcArgs a = bic_args(t, s->shape[im][fi], s->d_U, s->d_px, s->w, s->h, M);
                a.mode = 1;
                a.emitC = s->d_C;
                a.emitV = s->d_V;
                a.out = s->d_dummy;
                a.prev = s->d_dummy;
                a.frame0 = 0;
                a.nframes = s->nfmax;
                a.first_init = 1;
                a.rn = 1.0f;
                a.rn1 = 1.0f;
                a.sens = 1.0f;
                a.sens_int = 1;
                a.int_weighting = s->int_weighting;
                SCHK(s, launch_bic_rgc(a, s->shape[im][fi].ppw, st));
            }
            SweepAccArgs c{};
            c.C = s->d_C + s->cstride * slot;
            c.V = s->d_V;
            c.out = s->d_out[im];
            c.w = s->w;
            c.h = s->h;
            c.M = M;
            c.nframes = s->nfmax;
            c.nS = nS;
            c.n_fwhm = nF;
            c.i_fwhm = fi;
            c.nq = nQ;
            c.int_weighting = s->int_weighting;
            for (int q = 0; q < nQ; q++) {
                c.nf[q] = s->nfs[q];
                c.rn[q] = 1.0f / (float)s->nfs[q];
                c.rn1[q] = 1.0f / (float)(s->nfs[q] - 1);
            }
            for (int k = 0; k < nS; k++) {
                c.sens[k] = s->sens[k];
                c.sens_int[k] = sens_exact_int(s->sens[k]);
            }
            SCHK(s, launch_sweep_acc(c, st));
            // VAR and MP correction of each configuration (LiveSRRF_CL.java:753-770), the
            // nQ x nS configurations of this fwhm in one launch each: configuration
            // c = q nS + k sits at index c nF + fi
            const int nb = nQ * nS;
            float *out = s->d_out[im] + (long)fi * 4 * whM;
            SCHK(s, launch_var(out, whM, st, nb, (long)nF * 4 * whM));
            if (s->do_mp) {
                float *mp = s->d_mp[im] + (long)fi * 3 * M * M;
                SCHK(s, launch_mp_map(out, mp, s->w, s->h, 1, M, 1, st, nb, (long)nF * 4 * whM,
                                      (long)nF * 3 * M * M));
                SCHK(s, launch_mp_correct(out, mp, s->w, s->h, 1, M, 1, st, nb,
                                          (long)nF * 4 * whM, (long)nF * 3 * M * M));
            }
        }
    }
    return ESRRF_OK;
}

int tune(esrrf_sweep *s)
{
    if (s->multi) {
        // per (magnification, group of fwhm): the emit launch over up to 32 frames
        for (size_t im = 0; im < s->Ms.size(); im++) {
            const int M = s->Ms[im];
            for (size_t g = 0; g < s->mtabs[im].size(); g++) {
                const BicTablesMulti &t = s->mtabs[im][g];
                BicRgcArgs a = bic_args_multi(t, s->mshape[im][g], s->d_U, s->d_px, s->w, s->h,
                                              M);
                a.emitC = s->d_C;
                a.emitV = s->d_V;
                a.nframes = std::min(s->nfmax, 32);
                a.int_weighting = s->int_weighting;
                EmitShape best;
                SCHK(s, emit_autotune(t, a, M, s->cstride, 2, s->stream, &best));
                s->mshape[im][g] = best;
            }
        }
        return ESRRF_OK;
    }
    // per (magnification, fwhm): the emit launch over up to 32 frames
    for (size_t im = 0; im < s->Ms.size(); im++) {
        const int M = s->Ms[im];
        for (size_t fi = 0; fi < s->fwhm.size(); fi++) {
            const BicTables &t = s->tabs[im][fi];
            BicRgcArgs a = bic_args(t, s->shape[im][fi], s->d_U, s->d_px, s->w, s->h, M);
            a.mode = 1;
            a.emitC = s->d_C;
            a.emitV = s->d_V;
            a.out = s->d_dummy;
            a.prev = s->d_dummy;
            a.nframes = std::min(s->nfmax, 32);
            a.first_init = 1;
            a.rn = a.rn1 = a.sens = 1.0f;
            a.sens_int = 1;
            a.int_weighting = s->int_weighting;
            BicShape best;
            SCHK(s, bic_autotune(t, a, M, 2, s->stream, &best, nullptr));
            s->shape[im][fi] = best;
        }
    }
    return ESRRF_OK;
}

int run_impl(esrrf_sweep *s, const void *frames, int dtype, int n, bool on_device)
{
    if (dtype != ESRRF_U16 && dtype != ESRRF_F32) return sfail(s, ESRRF_ERR_INVALID, "dtype");
    if (n < s->nfmax)
        return sfail(s, ESRRF_ERR_INVALID, "the sweep needs %d frames, got %d", s->nfmax, n);
    (void)hipSetDevice(s->dev);
    const size_t bytes = (size_t)s->w * s->h * s->nfmax * (dtype == ESRRF_U16 ? 2 : 4);
    // first nfmax frames (gui/ParametersSweep_.java:808-816 takes frames 1..nf)
    SCHK(s, hipMemcpyAsync(s->d_in, frames, bytes,
                           on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s->stream));
    if (s->fht) s->tuned = true;   // FHT mode has one launch shape
    if (s->autotune && !s->tuned) {
        // tuning needs the gradients of these frames; it runs eagerly, before any capture
        const long wh = (long)s->w * s->h;
        SCHK(s, launch_convert_frames(s->d_in, dtype, s->d_px, wh * s->nfmax, s->stream));
        SCHK(s, launch_bic_prep(s->d_px, s->d_U, s->w, s->h, s->nfmax, s->stream));
        int rc = tune(s);
        if (rc) return rc;
        s->tuned = true;
    }
    if (!s->use_graph) {
        int rc = enqueue_run(s, dtype);
        if (rc) return rc;
        s->have_result = true;
        return ESRRF_OK;
    }
    if (!s->exec || s->graph_dtype != dtype) {
        if (s->exec) (void)hipGraphExecDestroy(s->exec);
        if (s->graph) (void)hipGraphDestroy(s->graph);
        s->exec = nullptr;
        s->graph = nullptr;
        SCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_run(s, dtype);
        hipGraph_t g = nullptr;
        hipError_t ec = hipStreamEndCapture(s->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess)
            return sfail(s, ESRRF_ERR_HIP, "graph capture: %s", hipGetErrorString(ec));
        s->graph = g;
        SCHK(s, hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0));
        s->graph_dtype = dtype;
    }
    SCHK(s, hipGraphLaunch(s->exec, s->stream));
    s->have_result = true;
    return ESRRF_OK;
}

int check_index(esrrf_sweep *s, int im, int iq, int is, int ifw)
{
    if (im < 0 || im >= (int)s->Ms.size() || iq < 0 || iq >= (int)s->nfs.size() || is < 0 ||
        is >= (int)s->sens.size() || ifw < 0 || ifw >= (int)s->fwhm.size())
        return sfail(s, ESRRF_ERR_INVALID, "configuration index out of range");
    if (!s->have_result) return sfail(s, ESRRF_ERR_STATE, "no sweep has run");
    return ESRRF_OK;
}

}  // namespace

int esrrf_sweep_create(const esrrf_sweep_params *p, esrrf_sweep **out)
{
    if (!p || !out) return ESRRF_ERR_INVALID;
    *out = nullptr;
    if (p->width < 4 || p->height < 4 || p->n_magnifications < 1 || p->n_fwhm < 1 ||
        p->n_sensitivity < 1 || p->n_n_frames < 1 || !p->magnifications || !p->fwhm ||
        !p->sensitivity || !p->n_frames || p->n_sensitivity > SWEEP_MAX_S ||
        p->n_n_frames > SWEEP_MAX_NF)
        return ESRRF_ERR_INVALID;
    esrrf_sweep *s = new esrrf_sweep();
    auto bad = [&](int code) {
        (void)hipGetLastError();   // a failed call must not surface in later checks
        sweep_free(s);
        delete s;
        return code;
    };
    s->w = p->width;
    s->h = p->height;
    s->Ms.assign(p->magnifications, p->magnifications + p->n_magnifications);
    s->fwhm.assign(p->fwhm, p->fwhm + p->n_fwhm);
    s->sens.assign(p->sensitivity, p->sensitivity + p->n_sensitivity);
    s->nfs.assign(p->n_frames, p->n_frames + p->n_n_frames);
    s->int_weighting = p->int_weighting ? 1 : 0;
    s->do_mp = p->do_mp_map_correction ? 1 : 0;
    s->autotune = p->autotune ? 1 : 0;
    s->use_graph = p->use_graph ? 1 : 0;
    s->multi = !kernel_variants_from_env().sweep_per_fwhm;
    s->fht = p->do_fht_interpolation ? 1 : 0;
    if (s->fht) s->multi = false;
    for (int M : s->Ms)
        if (M < 1 || M > 8) return bad(ESRRF_ERR_UNSUPPORTED);
    for (int nf : s->nfs) {
        if (nf < 1) return bad(ESRRF_ERR_INVALID);
        s->nfmax = std::max(s->nfmax, nf);
    }
    for (float f : s->fwhm)
        if (!(f > 0.0f)) return bad(ESRRF_ERR_INVALID);
    int dev = p->device;
    if (dev < 0 && esrrf_fastest_device(&dev) != ESRRF_OK) return bad(ESRRF_ERR_HIP);
    s->dev = dev;
    if (hipSetDevice(dev) != hipSuccess) return bad(ESRRF_ERR_HIP);
    if (hipStreamCreate(&s->own_stream) != hipSuccess) return bad(ESRRF_ERR_HIP);
    s->stream = s->own_stream;
    const long wh = (long)s->w * s->h;
    int MMmax = 1;
    for (int M : s->Ms) MMmax = std::max(MMmax, M * M);
    s->cstride = wh * MMmax * s->nfmax;
    auto alloc = [&](void **ptr, size_t bytes) {
        return hipMalloc(ptr, bytes ? bytes : 4) == hipSuccess;
    };
    if (!alloc(&s->d_in, sizeof(float) * wh * s->nfmax) ||
        !alloc((void **)&s->d_px, sizeof(float) * wh * s->nfmax) ||
        !alloc((void **)&s->d_U, sizeof(float2) * 4 * wh * s->nfmax) ||
        !alloc((void **)&s->d_C, sizeof(float) * wh * MMmax * s->nfmax *
                                     (s->multi ? std::min((int)s->fwhm.size(), SWEEP_NFG) : 1)) ||
        !alloc((void **)&s->d_V, sizeof(float) * wh * MMmax * s->nfmax) ||
        !alloc((void **)&s->d_dummy, 64))
        return bad(ESRRF_ERR_NOMEM);
    s->tabs.resize(s->Ms.size());
    s->shape.resize(s->Ms.size());
    s->d_out.assign(s->Ms.size(), nullptr);
    s->d_mp.assign(s->Ms.size(), nullptr);
    const size_t nc = n_conf(s);
    if (s->fht) {
        const KernelVariants var = kernel_variants_from_env();
        s->fms.resize(s->Ms.size());
        s->f_px.assign(s->Ms.size(), nullptr);
        s->f_U.assign(s->Ms.size(), nullptr);
        s->f_aux.assign(s->Ms.size(), nullptr);
        s->ftaps.resize(s->Ms.size());
        for (size_t im = 0; im < s->Ms.size(); im++) {
            const int M = s->Ms[im];
            float sig = 0, rad = 0, tss = 0, tso = 0;
            esrrf_fht_constants(s->fwhm[0], M, &sig, &rad, &tss, &tso);
            ModeCreate mc{};
            mc.mode = 1;
            mc.w = mc.W = s->w;
            mc.h = mc.H = s->h;
            mc.nP = 1;
            mc.M = M;
            mc.nF = s->nfmax;
            mc.radius = rad;
            mc.tSS = tss;
            mc.tSO = tso;
            mc.var = var;
            size_t auxb = 0;
            int rc = mode_create(mc, s->fms[im], &s->f_px[im], &s->f_U[im], &s->f_aux[im], &auxb);
            if (rc != ESRRF_OK) {
                sfail(s, rc, "FHT state: %s", s->fms[im].err.c_str());
                return bad(rc);
            }
            s->ftaps[im].resize(s->fwhm.size());
            for (size_t fi = 0; fi < s->fwhm.size(); fi++) {
                esrrf_fht_constants(s->fwhm[fi], M, &sig, &rad, &tss, &tso);
                rc = fht_taps_create(rad, tss, tso, s->ftaps[im][fi]);
                if (rc != ESRRF_OK) return bad(rc);
                const FhtTaps &t = s->ftaps[im][fi];
                // the emit launch keeps the window's tile and the taps in LDS
                if (fht_rgc_lds_bytes(t.TW, t.pitch, t.ntaps) > 65536)
                    return bad(ESRRF_ERR_UNSUPPORTED);
            }
        }
    }
    for (size_t im = 0; im < s->Ms.size(); im++) {
        const int M = s->Ms[im];
        if (s->fht) {
            if (!alloc((void **)&s->d_out[im], sizeof(float) * nc * 4 * wh * M * M) ||
                !alloc((void **)&s->d_mp[im], sizeof(float) * nc * 3 * M * M))
                return bad(ESRRF_ERR_NOMEM);
            s->tmp_floats = std::max(s->tmp_floats, (size_t)4 * wh * M * M);
            continue;
        }
        s->tabs[im].resize(s->fwhm.size());
        for (size_t fi = 0; fi < s->fwhm.size(); fi++) {
            if (bic_tables_create(s->w, s->h, M, s->fwhm[fi], s->tabs[im][fi]) != hipSuccess)
                return bad(ESRRF_ERR_HIP);
            if (!s->tabs[im][fi].ok) return bad(ESRRF_ERR_UNSUPPORTED);
        }
        s->shape[im].resize(s->fwhm.size());
        for (size_t fi = 0; fi < s->fwhm.size(); fi++)
            s->shape[im][fi] = bic_default_shape(s->tabs[im][fi], M);
        if (s->multi) {
            const int ng = ((int)s->fwhm.size() + SWEEP_NFG - 1) / SWEEP_NFG;
            s->mtabs.resize(s->Ms.size());
            s->mshape.resize(s->Ms.size());
            s->mtabs[im].resize(ng);
            s->mshape[im].resize(ng);
            for (int g = 0; g < ng; g++) {
                const int f0 = g * SWEEP_NFG;
                const int nfg = std::min(SWEEP_NFG, (int)s->fwhm.size() - f0);
                if (bic_multi_tables_create(s->w, s->h, M, s->fwhm.data() + f0, nfg,
                                            s->mtabs[im][g]) != hipSuccess)
                    return bad(ESRRF_ERR_HIP);
                if (!s->mtabs[im][g].ok) return bad(ESRRF_ERR_UNSUPPORTED);
                s->mshape[im][g] = emit_candidate_shapes(s->mtabs[im][g], M)[0];
            }
        }
        if (!alloc((void **)&s->d_out[im], sizeof(float) * nc * 4 * wh * M * M) ||
            !alloc((void **)&s->d_mp[im], sizeof(float) * nc * 3 * M * M))
            return bad(ESRRF_ERR_NOMEM);
        s->tmp_floats = std::max(s->tmp_floats, (size_t)4 * wh * M * M);
    }
    if (!alloc((void **)&s->d_tmp, sizeof(float) * s->tmp_floats)) return bad(ESRRF_ERR_NOMEM);
    *out = s;
    return ESRRF_OK;
}

int esrrf_sweep_destroy(esrrf_sweep *s)
{
    if (!s) return ESRRF_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        (void)hipSetDevice(s->dev);
        (void)hipStreamSynchronize(s->stream);
        sweep_free(s);
    }
    delete s;
    return ESRRF_OK;
}

int esrrf_sweep_set_stream(esrrf_sweep *s, void *hip_stream)
{
    if (!s) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    s->stream = hip_stream ? (hipStream_t)hip_stream : s->own_stream;
    return ESRRF_OK;
}

int esrrf_sweep_run(esrrf_sweep *s, const void *frames, int dtype, int n_frames)
{
    if (!s || !frames) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    int rc = run_impl(s, frames, dtype, n_frames, false);
    if (rc) return rc;
    SCHK(s, hipStreamSynchronize(s->stream));
    return ESRRF_OK;
}

int esrrf_sweep_run_device(esrrf_sweep *s, const void *d_frames, int dtype, int n_frames)
{
    if (!s || !d_frames) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    return run_impl(s, d_frames, dtype, n_frames, true);
}

int esrrf_sweep_read(esrrf_sweep *s, int i_mag, int i_nf, int i_sens, int i_fwhm, float *out)
{
    if (!s || !out) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    int rc = check_index(s, i_mag, i_nf, i_sens, i_fwhm);
    if (rc) return rc;
    (void)hipSetDevice(s->dev);
    const int M = s->Ms[i_mag];
    const long n = 4L * s->w * s->h * M * M;
    const long ci = ((long)i_nf * s->sens.size() + i_sens) * s->fwhm.size() + i_fwhm;
    SCHK(s, launch_nan_to_zero_copy(s->d_out[i_mag] + ci * n, s->d_tmp, n, s->stream));
    SCHK(s, hipMemcpyAsync(out, s->d_tmp, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream));
    SCHK(s, hipStreamSynchronize(s->stream));
    return ESRRF_OK;
}

int esrrf_sweep_read_device(esrrf_sweep *s, int i_mag, int i_nf, int i_sens, int i_fwhm,
                            float *d_out)
{
    if (!s || !d_out) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    int rc = check_index(s, i_mag, i_nf, i_sens, i_fwhm);
    if (rc) return rc;
    (void)hipSetDevice(s->dev);
    const int M = s->Ms[i_mag];
    const long n = 4L * s->w * s->h * M * M;
    const long ci = ((long)i_nf * s->sens.size() + i_sens) * s->fwhm.size() + i_fwhm;
    SCHK(s, launch_nan_to_zero_copy(s->d_out[i_mag] + ci * n, d_out, n, s->stream));
    return ESRRF_OK;
}

int esrrf_sweep_shape(esrrf_sweep *s, int i_mag, int i_fwhm, int *nw, int *ppw, int *taps_in_lds,
                      int *frames_per_pass, float *ms)
{
    if (!s || i_mag < 0 || i_mag >= (int)s->Ms.size() || i_fwhm < 0 ||
        i_fwhm >= (int)s->fwhm.size())
        return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fht) return sfail(s, ESRRF_ERR_STATE, "FHT mode has no tiled 2-D launch shape");
    if (s->multi) {
        const EmitShape &m = s->mshape[i_mag][i_fwhm / SWEEP_NFG];
        if (nw) *nw = m.nw;
        if (ppw) *ppw = m.ppw;
        if (taps_in_lds) *taps_in_lds = 1;
        if (frames_per_pass) *frames_per_pass = 2;
        if (ms) *ms = m.ms;
        return ESRRF_OK;
    }
    const BicShape &b = s->shape[i_mag][i_fwhm];
    if (nw) *nw = b.nw;
    if (ppw) *ppw = b.ppw;
    if (taps_in_lds) *taps_in_lds = b.taps_in_lds;
    if (frames_per_pass) *frames_per_pass = b.fp;
    if (ms) *ms = b.ms;
    return ESRRF_OK;
}

const char *esrrf_sweep_last_error(esrrf_sweep *s) { return s ? s->err.c_str() : "null sweep"; }

int esrrf_sweep_squirrel(esrrf_sweep *s, int channel, int crop_size, int fix_sigma,
                         float fixed_sigma, double *rse, double *rsp, double *sigma)
{
    if (!s || channel < 0 || channel > 2) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->have_result) return sfail(s, ESRRF_ERR_STATE, "no sweep has run");
    if (crop_size < 0 || 2 * crop_size >= std::min(s->w, s->h))
        return sfail(s, ESRRF_ERR_INVALID, "crop_size %d outside [0, %d)", crop_size,
                     (std::min(s->w, s->h) + 1) / 2);
    SCHK(s, hipSetDevice(s->dev));
    // the configurations run on SQ_WORKERS host threads, each with its own SQUIRREL plan and
    // stream: a Brent search is a sequence of short launches and host reads, so searches of
    // different configurations overlap on the device; each writes only its own results
    SCHK(s, hipStreamSynchronize(s->stream));   // the sweep's outputs are complete
    constexpr int SQ_WORKERS = 8;
    const long ncfg = (long)s->nfs.size() * s->sens.size() * s->fwhm.size();
    for (size_t im = 0; im < s->Ms.size(); im++) {
        const int M = s->Ms[im];
        esrrf_squirrel_params qp{};
        qp.width = s->w;
        qp.height = s->h;
        qp.magnification = M;
        qp.crop_border = 1;
        qp.crop_size = crop_size;
        qp.fix_sigma = fix_sigma;
        qp.fixed_sigma = fixed_sigma;
        qp.device = s->dev;
        const int nw = (int)std::min<long>(SQ_WORKERS, ncfg);
        std::vector<esrrf_squirrel *> qs(nw, nullptr);
        int rc = ESRRF_OK;
        for (int t = 0; t < nw && rc == ESRRF_OK; t++) {
            rc = esrrf_squirrel_create(&qp, &qs[t]);
            if (rc) s->err = "SQUIRREL plan for M=" + std::to_string(M) + " failed";
        }
        const long whM = (long)s->w * s->h * M * M;
        std::vector<int> rcs(nw, ESRRF_OK);
        std::vector<std::string> errs(nw);
        auto work = [&](int t) {
            for (long c = t; c < ncfg; c += nw) {
                // configuration c of [nf][S][fwhm]: SR = its channel, ref = its INT (:473)
                const float *base = s->d_out[im] + c * 4 * whM;
                esrrf_squirrel_result r{};
                const int e = esrrf_squirrel_run_device(qs[t], base + 3 * whM,
                                                        base + channel * whM, &r, nullptr,
                                                        nullptr);
                if (e) {
                    rcs[t] = e;
                    errs[t] = std::string("SQUIRREL: ") + esrrf_squirrel_last_error(qs[t]);
                    return;
                }
                const long o = (long)im * ncfg + c;
                if (rse) rse[o] = r.rse;
                if (rsp) rsp[o] = r.rsp;
                if (sigma) sigma[o] = r.sigma;
            }
        };
        if (rc == ESRRF_OK) {
            // a worker the system cannot start is run on this thread instead (no exception
            // leaves the C ABI)
            std::vector<std::thread> th;
            int started = 1;
            try {
                for (; started < nw; started++) th.emplace_back(work, started);
            } catch (...) {
            }
            work(0);
            for (int t = started; t < nw; t++) work(t);
            for (auto &x : th) x.join();
            for (int t = 0; t < nw && rc == ESRRF_OK; t++)
                if (rcs[t]) {
                    rc = rcs[t];
                    s->err = errs[t];
                }
        }
        for (esrrf_squirrel *q : qs)
            if (q) esrrf_squirrel_destroy(q);
        if (rc) return rc;
    }
    return ESRRF_OK;
}

int esrrf_sweep_frc(esrrf_sweep *odd, esrrf_sweep *even, int channel, double *resolution)
{
    if (!odd || !even || odd == even || channel < 0 || channel > 2 || !resolution)
        return ESRRF_ERR_INVALID;
    std::scoped_lock lk(odd->mu, even->mu);
    if (!odd->have_result || !even->have_result)
        return sfail(odd, ESRRF_ERR_STATE, "both sweeps must have run");
    if (odd->Ms != even->Ms || odd->fwhm != even->fwhm || odd->sens != even->sens ||
        odd->nfs != even->nfs || odd->w != even->w || odd->h != even->h || odd->dev != even->dev)
        return sfail(odd, ESRRF_ERR_INVALID, "the two sweeps must have the same parameters");
    SCHK(odd, hipSetDevice(odd->dev));
    SCHK(odd, hipStreamSynchronize(even->stream));
    const long ncfg = (long)odd->nfs.size() * odd->sens.size() * odd->fwhm.size();
    for (size_t im = 0; im < odd->Ms.size(); im++) {
        const int M = odd->Ms[im];
        esrrf_frc *q = nullptr;
        int rc = esrrf_frc_create(odd->w * M, odd->h * M, odd->dev, &q);
        if (rc) return sfail(odd, rc, "FRC plan for M=%d: error %d", M, rc);
        esrrf_frc_set_stream(q, odd->stream);
        const long whM = (long)odd->w * odd->h * M * M;
        for (long c = 0; c < ncfg && rc == ESRRF_OK; c++) {
            const long off = c * 4 * whM + channel * whM;
            rc = esrrf_frc_run_device(q, odd->d_out[im] + off, even->d_out[im] + off,
                                      &resolution[(long)im * ncfg + c], nullptr);
            if (rc) odd->err = std::string("FRC: ") + esrrf_frc_last_error(q);
        }
        esrrf_frc_destroy(q);
        if (rc) return rc;
    }
    return ESRRF_OK;
}
