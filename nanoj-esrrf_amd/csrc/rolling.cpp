// rolling.cpp -- rolling analysis of a long stack (gui/LiveSRRF_optimised_.java:276-384):
// SR image r covers frames [r gap, r gap + nF) (:621-627, :842-877), 2-D BIC, no drift.
//
// The frames stream through in chunks: per chunk one conversion, one gradient pass and one
// RGC launch in its emit variant (per-frame normalised CGLH and intensity), then one
// accumulation launch that adds the chunk's frames to every SR window they belong to, each
// at its own index inside the window (k_rolling.hip).  Overlapping windows share the RGC work
// that the plugin repeats per window.  A window whose last frame has been accumulated gets
// VAR and the MP correction (LiveSRRF_CL.java:753-770) and waits in its slot until
// esrrf_rolling_next hands it out, in order; the caller streams it to disk or keeps it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/esrrf.h"
#include "bic_launch.h"
#include "esrrf_internal.h"
#include "rolling_internal.h"

using namespace esrrf;

struct esrrf_rolling {
    std::mutex mu;
    std::string err;
    int dev = 0;
    hipStream_t own_stream = nullptr, stream = nullptr;
    int w = 0, h = 0, M = 1;
    float fwhm = 0.0f, sens = 1.0f;
    int sens_int = 1, intw = 1, mp = 1;
    int nF_user = 0, gap_user = 0, chunk_user = 256;
    BicTables tabs;
    BicShape shape;
    // run
    const char *src = nullptr;
    bool src_device = false;
    int dtype = 0;
    int nS = 0, nF = 0, gap = 0, nSR = 0;
    int chunk = 0;            // frames per chunk for this run
    int f_done = 0;           // frames [0, f_done) accumulated
    int next_out = 0;         // next SR image to hand out
    int n_complete = 0;       // SR images [0, n_complete) finalised
    bool running = false;
    // device
    int nslots = 0;
    size_t chunk_cap = 0;     // frames the chunk buffers hold
    void *d_in = nullptr;     // host-stream staging (raw dtype)
    float *d_px = nullptr;
    float2 *d_U = nullptr;
    float *d_C = nullptr, *d_V = nullptr, *d_dummy = nullptr;
    float *d_state = nullptr; // [nslots][5][M*M][h][w]; [4][hM][wM] once complete
    float *d_mp = nullptr;
    float *d_tmp = nullptr;   // image-order and read-out staging [4][hM][wM]
};

namespace {

int lfail(esrrf_rolling *r, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (r) r->err = buf;
    return code;
}

#define LCHK(r, x)                                                                         \
    do {                                                                                   \
        hipError_t _st = (x);                                                              \
        if (_st != hipSuccess)                                                             \
            return lfail((r), ESRRF_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(_st)); \
    } while (0)

void free_run_buffers(esrrf_rolling *r)
{
    void *ptrs[] = {r->d_in, r->d_px, r->d_U, r->d_C, r->d_V, r->d_state, r->d_tmp};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    r->d_in = r->d_px = nullptr;
    r->d_U = nullptr;
    r->d_C = r->d_V = r->d_state = r->d_tmp = nullptr;
    r->chunk_cap = 0;
    r->nslots = 0;
}

void rolling_free(esrrf_rolling *r)
{
    free_run_buffers(r);
    if (r->d_dummy) (void)hipFree(r->d_dummy);
    if (r->d_mp) (void)hipFree(r->d_mp);
    bic_tables_free(r->tabs);
    if (r->own_stream) (void)hipStreamDestroy(r->own_stream);
}

long whM_of(const esrrf_rolling *r) { return (long)r->w * r->h * r->M * r->M; }

float *slot_of(esrrf_rolling *r, int win)
{
    return r->d_state + (size_t)(win % r->nslots) * 5 * whM_of(r);
}

int begin_impl(esrrf_rolling *r, const void *frames, bool on_device, int dtype, int n_slices,
               int *n_sr)
{
    if (!frames || (dtype != ESRRF_U16 && dtype != ESRRF_F32) || n_slices < 2)
        return lfail(r, ESRRF_ERR_INVALID, "frames, dtype and n_slices >= 2 required");
    // gui/LiveSRRF_optimised_.java:621-627 (rolling analysis on)
    int plan[6];
    int rc = esrrf_plan_rolling(n_slices, r->nF_user, r->gap_user, 1, n_slices, plan);
    if (rc) return lfail(r, rc, "rolling plan of %d slices", n_slices);
    const int nF = plan[0], gap = plan[1], nSR = plan[2];
    if (nF < 2) return lfail(r, ESRRF_ERR_INVALID, "an SR image needs at least 2 frames");
    // windows overlapping one chunk of n frames: at most ceil((n + nF) / gap) + 1
    const int maxwin = ROLL_MAX_WIN;
    long n = std::min<long>(r->chunk_user, (long)(maxwin - 2) * gap - nF);
    if (n < 1)
        return lfail(r, ESRRF_ERR_UNSUPPORTED,
                     "frame gap %d too small for %d frames per SR image (at most %d windows "
                     "may overlap a frame)", gap, nF, maxwin - 2);
    n = std::min<long>(n, (long)(nSR - 1) * gap + nF);
    const int nslots = (int)std::min<long>(maxwin, (n + nF + gap - 1) / gap + 1);
    const long wh = (long)r->w * r->h, whM = whM_of(r);
    LCHK(r, hipSetDevice(r->dev));
    if (r->chunk_cap < (size_t)n || r->nslots != nslots) {
        free_run_buffers(r);
        auto alloc = [&](void **p, size_t b) { return hipMalloc(p, b ? b : 8) == hipSuccess; };
        if (!alloc(&r->d_in, sizeof(float) * wh * n) ||
            !alloc((void **)&r->d_px, sizeof(float) * wh * n) ||
            !alloc((void **)&r->d_U, sizeof(float2) * 4 * wh * n) ||
            !alloc((void **)&r->d_C, sizeof(float) * whM * n) ||
            !alloc((void **)&r->d_V, sizeof(float) * whM * n) ||
            !alloc((void **)&r->d_state, sizeof(float) * 5 * whM * nslots) ||
            !alloc((void **)&r->d_tmp, sizeof(float) * 4 * whM)) {
            free_run_buffers(r);
            return lfail(r, ESRRF_ERR_NOMEM, "rolling buffers (%ld frames per chunk, %d slots)",
                         n, nslots);
        }
        r->chunk_cap = (size_t)n;
        r->nslots = nslots;
    }
    r->src = static_cast<const char *>(frames);
    r->src_device = on_device;
    r->dtype = dtype;
    r->nS = n_slices;
    r->nF = nF;
    r->gap = gap;
    r->nSR = nSR;
    r->chunk = (int)n;
    r->f_done = 0;
    r->next_out = 0;
    r->n_complete = 0;
    r->running = true;
    if (n_sr) *n_sr = nSR;
    return ESRRF_OK;
}

// Accumulate the next chunk of frames into every window it touches; finalise the windows
// it completes.
int step_chunk(esrrf_rolling *r)
{
    const hipStream_t st = r->stream;
    const long wh = (long)r->w * r->h, whM = whM_of(r);
    const int nF = r->nF, gap = r->gap;
    // skip frames that belong to no window (frame gap > frames per SR image)
    int c0 = r->f_done;
    const int first_open = r->n_complete;   // windows [first_open, nSR) are unfinished
    if (first_open >= r->nSR) return lfail(r, ESRRF_ERR_STATE, "every SR image is complete");
    c0 = std::max(c0, first_open * gap);
    int c1 = std::min(c0 + r->chunk, (r->nSR - 1) * gap + nF);
    // end of the frames any window starting before c1 needs
    {
        const int last_start = std::min(r->nSR - 1, (c1 - 1) / gap);
        c1 = std::min(c1, last_start * gap + nF);
    }
    const int n = c1 - c0;
    if (n <= 0) return lfail(r, ESRRF_ERR_STATE, "empty chunk");
    const size_t esz = r->dtype == ESRRF_U16 ? 2 : 4;
    const void *src = r->src + (size_t)c0 * wh * esz;
    if (!r->src_device) {
        LCHK(r, hipMemcpyAsync(r->d_in, src, esz * wh * n, hipMemcpyHostToDevice, st));
        src = r->d_in;
    }
    LCHK(r, launch_convert_frames(src, r->dtype, r->d_px, wh * n, st));
    LCHK(r, launch_bic_prep(r->d_px, r->d_U, r->w, r->h, n, st));
    BicRgcArgs a = bic_args(r->tabs, r->shape, r->d_U, r->d_px, r->w, r->h, r->M);
    a.mode = 1;
    a.emitC = r->d_C;
    a.emitV = r->d_V;
    a.out = r->d_dummy;
    a.prev = r->d_dummy;
    a.frame0 = 0;
    a.nframes = n;
    a.first_init = 1;
    a.rn = a.rn1 = 1.0f;
    a.sens = 1.0f;
    a.sens_int = 1;
    a.int_weighting = r->intw;
    LCHK(r, launch_bic_rgc(a, r->shape.ppw, st));
    // windows overlapping [c0, c1): r gap < c1 and r gap + nF > c0 -- consecutive windows
    // from the first unfinished one, accumulated in launches of up to ROLL_GROUP
    int wlo = first_open;
    while (wlo < r->nSR && wlo * gap + nF <= c0) wlo++;
    int whi = wlo;   // one past the last overlapping window
    while (whi < r->nSR && whi * gap < c1) whi++;
    if (whi - wlo > ROLL_MAX_WIN) return lfail(r, ESRRF_ERR_STATE, "too many windows");
    if (whi - 1 - r->next_out >= r->nslots)
        return lfail(r, ESRRF_ERR_STATE, "window %d has no free slot", whi - 1);
    RollAccArgs c{};
    c.C = r->d_C;
    c.V = r->d_V;
    c.state = r->d_state;
    c.w = r->w;
    c.h = r->h;
    c.M = r->M;
    c.nframes = n;
    c.nF = nF;
    c.c0 = c0;
    c.gap = gap;
    c.nslots = r->nslots;
    c.sens = r->sens;
    c.sens_int = r->sens_int;
    c.int_weighting = r->intw;
    c.rn = 1.0f / (float)nF;
    c.rn1 = 1.0f / (float)(nF - 1);
    for (int g = wlo; g < whi; g += ROLL_GROUP) {
        c.r0 = g;
        c.nwin = std::min(ROLL_GROUP, whi - g);
        LCHK(r, launch_roll_acc(c, st));
    }
    std::vector<int> done;
    for (int win = wlo; win < whi; win++)
        if (win * gap + nF <= c1) done.push_back(win);
    for (int win : done) {
        float *out = slot_of(r, win);
        LCHK(r, launch_roll_to_image(out, r->d_tmp, r->w, r->h, r->M, st));
        LCHK(r, hipMemcpyAsync(out, r->d_tmp, sizeof(float) * 4 * whM, hipMemcpyDeviceToDevice,
                               st));
        LCHK(r, launch_var(out, whM, st));
        if (r->mp) {
            LCHK(r, launch_mp_map(out, r->d_mp, r->w, r->h, 1, r->M, 1, st));
            LCHK(r, launch_mp_correct(out, r->d_mp, r->w, r->h, 1, r->M, 1, st));
        }
        r->n_complete = std::max(r->n_complete, win + 1);
    }
    r->f_done = c1;
    return ESRRF_OK;
}

int next_impl(esrrf_rolling *r, float *out, bool to_device, int *index)
{
    if (!r->running) return lfail(r, ESRRF_ERR_STATE, "call esrrf_rolling_begin first");
    if (r->next_out >= r->nSR) return lfail(r, ESRRF_ERR_STATE, "every SR image was read");
    LCHK(r, hipSetDevice(r->dev));
    while (r->n_complete <= r->next_out) {
        const int rc = step_chunk(r);
        if (rc) return rc;
    }
    const long whM = whM_of(r);
    // readSRRFbuffer's NaN -> 0 (LiveSRRF_CL.java:776-781)
    float *dst = to_device ? out : r->d_tmp;
    LCHK(r, launch_nan_to_zero_copy(slot_of(r, r->next_out), dst, 4 * whM, r->stream));
    if (!to_device) {
        LCHK(r, hipMemcpyAsync(out, r->d_tmp, sizeof(float) * 4 * whM, hipMemcpyDeviceToHost,
                               r->stream));
        LCHK(r, hipStreamSynchronize(r->stream));
    }
    if (index) *index = r->next_out;
    r->next_out++;
    return ESRRF_OK;
}

}  // namespace

int esrrf_rolling_create(const esrrf_rolling_params *p, esrrf_rolling **out)
{
    if (!p || !out) return ESRRF_ERR_INVALID;
    *out = nullptr;
    if (p->width < 4 || p->height < 4 || p->magnification < 1 || p->magnification > 8 ||
        !(p->fwhm > 0.0f) || !(p->sensitivity > 0.0f) || p->n_frame_for_srrf < 0 ||
        p->frame_gap < 0)
        return ESRRF_ERR_INVALID;
    esrrf_rolling *r = new esrrf_rolling();
    auto bad = [&](int code) {
        (void)hipGetLastError();   // a failed call must not surface in later checks
        rolling_free(r);
        delete r;
        return code;
    };
    r->w = p->width;
    r->h = p->height;
    r->M = p->magnification;
    r->fwhm = p->fwhm;
    r->sens = p->sensitivity;
    r->sens_int = sens_exact_int(p->sensitivity);
    r->intw = p->int_weighting ? 1 : 0;
    r->mp = p->do_mp_map_correction ? 1 : 0;
    r->nF_user = p->n_frame_for_srrf;
    r->gap_user = p->frame_gap;
    r->chunk_user = p->chunk_frames > 0 ? p->chunk_frames : 256;
    int dev = p->device;
    if (dev < 0 && esrrf_fastest_device(&dev) != ESRRF_OK) return bad(ESRRF_ERR_HIP);
    r->dev = dev;
    if (hipSetDevice(dev) != hipSuccess) return bad(ESRRF_ERR_HIP);
    if (hipStreamCreate(&r->own_stream) != hipSuccess) return bad(ESRRF_ERR_HIP);
    r->stream = r->own_stream;
    if (bic_tables_create(r->w, r->h, r->M, r->fwhm, r->tabs) != hipSuccess)
        return bad(ESRRF_ERR_HIP);
    if (!r->tabs.ok) return bad(ESRRF_ERR_UNSUPPORTED);
    r->shape = bic_default_shape(r->tabs, r->M, false);
    if (hipMalloc((void **)&r->d_dummy, 64) != hipSuccess ||
        hipMalloc((void **)&r->d_mp, sizeof(float) * 3 * r->M * r->M) != hipSuccess)
        return bad(ESRRF_ERR_NOMEM);
    *out = r;
    return ESRRF_OK;
}

int esrrf_rolling_destroy(esrrf_rolling *r)
{
    if (!r) return ESRRF_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(r->mu);
        (void)hipSetDevice(r->dev);
        (void)hipStreamSynchronize(r->stream);
        rolling_free(r);
    }
    delete r;
    return ESRRF_OK;
}

int esrrf_rolling_set_stream(esrrf_rolling *r, void *hip_stream)
{
    if (!r) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(r->mu);
    r->stream = hip_stream ? (hipStream_t)hip_stream : r->own_stream;
    return ESRRF_OK;
}

int esrrf_rolling_begin(esrrf_rolling *r, const void *frames, int dtype, int n_slices, int *n_sr)
{
    if (!r) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(r->mu);
    return begin_impl(r, frames, false, dtype, n_slices, n_sr);
}

int esrrf_rolling_begin_device(esrrf_rolling *r, const void *d_frames, int dtype, int n_slices,
                               int *n_sr)
{
    if (!r) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(r->mu);
    return begin_impl(r, d_frames, true, dtype, n_slices, n_sr);
}

int esrrf_rolling_next(esrrf_rolling *r, float *out, int *index)
{
    if (!r || !out) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(r->mu);
    return next_impl(r, out, false, index);
}

int esrrf_rolling_next_device(esrrf_rolling *r, float *d_out, int *index)
{
    if (!r || !d_out) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(r->mu);
    return next_impl(r, d_out, true, index);
}

const char *esrrf_rolling_last_error(esrrf_rolling *r) { return r ? r->err.c_str() : ""; }
