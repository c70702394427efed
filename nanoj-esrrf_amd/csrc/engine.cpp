// engine.cpp -- the eSRRF engine behind the C ABI of include/esrrf.h.
//
// It owns what LiveSRRF_CL owns in the reference (LiveSRRF_CL.java:29-102): the device
// buffers, the derived constants, the two frame counters, and the enqueue order of
// calculateSRRF / readSRRFbuffer.  The counters live on the host (exact integers, returned
// by esrrf_read_frame_counters) and are passed to the kernels as arguments, instead of the
// reference's device-side int[2] bumped by a kernel launch per frame
// (liveSRRF.cl:439-461, LiveSRRF_CL.java:737-740).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/esrrf.h"
#include "bic_launch.h"
#include "esrrf_internal.h"
#include "geometry.h"
#include "modes.h"
#include "variants.h"

using namespace esrrf;

namespace {

enum Mode { MODE_BIC = 0, MODE_FHT = 1, MODE_3D = 2 };

// Most upload chunks one load of host frames is split into.
constexpr int kMaxUploadChunks = 16;

struct ProfSlot {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double total_ms = 0.0;
    long launches = 0;
};

}  // namespace

struct esrrf_engine {
    std::mutex mu;
    std::string err;
    int dev = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    // host-frame upload overlapped with compute: chunks go up on copy_stream, each with an
    // event the compute stream waits for (BIC mode)
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_free = nullptr;                 // staging buffer free to overwrite
    std::vector<hipEvent_t> ev_chunk;             // chunk c resident
    hipEvent_t ev_int = nullptr;                  // read-back: INT channel ready to copy
    int upload_chunks = 4;
    Mode mode = MODE_BIC;
    esrrf_params p{};
    // sizes
    int W = 0, H = 0;            // raw frame (3-D: montage)
    int w = 0, h = 0, nP = 1;    // per plane (widthS, heightS, nPlanes)
    int M = 1, wM = 0, hM = 0, nPlanesM = 1;
    int Mz = 1;                  // axial magnification (3-D), 1 in 2-D
    long nvox = 0;               // voxels of one output channel
    float sigma = 0, radius = 0, tSS = 0, tSO = 0;
    int sens_int = 0;
    // device memory
    void *d_in = nullptr;        // staging of host frames (raw dtype)
    size_t d_in_bytes = 0;
    float *d_px = nullptr;       // [nF][planes][h][w] float (FHT: magnified frames)
    float2 *d_U = nullptr;       // unit tap gradients of the load
    float *d_out = nullptr;      // [4][nvox]
    float *d_prev = nullptr;     // [nvox]
    float *d_mp = nullptr;       // MP map
    float *d_drift = nullptr;    // [2][nFrameForSRRF]
    float *d_tmp = nullptr;      // read-back staging [4][nvox]
    float *d_aux = nullptr;      // mode scratch (FHT spectra, 3-D gradients)
    size_t d_aux_bytes = 0;
    BicTables tabs;              // tap tables of the tiled BIC kernel
    BicShape shape;              // its launch shape (default or autotuned)
    bool tiled = false;          // tiled BIC kernel usable
    int stage = 0;
    int pair_direct = 0;
    KernelVariants var;          // alternative kernels (read once at create)
    std::vector<float> h_drift;  // host copy of the drift buffer (x M)
    ModeState ms;                // FHT / 3-D state
    // bookkeeping
    int counters[2] = {0, 0};
    int loaded = 0;              // frames of the current load
    bool has_mp = false;
    bool profiling = false;
    ProfSlot prof[3];
    const char *rgc_kernel = "";   // RGC kernel of the last launch (BIC mode)
};

namespace {

int fail(esrrf_engine *e, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (e) e->err = buf;
    return code;
}

#define HIPCHK(e, x)                                                                     \
    do {                                                                                 \
        hipError_t _st = (x);                                                            \
        if (_st != hipSuccess)                                                           \
            return fail((e), ESRRF_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(_st)); \
    } while (0)

template <class T>
int dalloc(esrrf_engine *e, T **ptr, size_t bytes)
{
    *ptr = nullptr;
    if (bytes == 0) bytes = 4;
    hipError_t st = hipMalloc((void **)ptr, bytes);
    if (st != hipSuccess) {
        *ptr = nullptr;
        (void)hipGetLastError();   // not to surface in the next launch's error check
        return fail(e, ESRRF_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(st));
    }
    return ESRRF_OK;
}

void free_all(esrrf_engine *e)
{
    void *ptrs[] = {e->d_in, e->d_px, e->d_U, e->d_out, e->d_prev, e->d_mp, e->d_drift,
                    e->d_tmp, e->d_aux};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    free_mode_state(e->ms);
    bic_tables_free(e->tabs);
    for (auto &s : e->prof)
        for (auto &ev : s.pending) {
            (void)hipEventDestroy(ev.first);
            (void)hipEventDestroy(ev.second);
        }
    for (hipEvent_t ev : e->ev_chunk)
        if (ev) (void)hipEventDestroy(ev);
    if (e->ev_free) (void)hipEventDestroy(e->ev_free);
    if (e->ev_int) (void)hipEventDestroy(e->ev_int);
    if (e->copy_stream) (void)hipStreamDestroy(e->copy_stream);
    if (e->own_stream) (void)hipStreamDestroy(e->own_stream);
}


// Record an event pair around a launch when profiling (HIP events on the launch stream).
struct ProfScope {
    esrrf_engine *e;
    int slot;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(esrrf_engine *e_, int s) : e(e_), slot(s)
    {
        if (!e->profiling) return;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            a = b = nullptr;
            return;
        }
        (void)hipEventRecord(a, e->stream);
    }
    ~ProfScope()
    {
        if (!a) return;
        (void)hipEventRecord(b, e->stream);
        e->prof[slot].pending.emplace_back(a, b);
    }
};

void prof_collect(esrrf_engine *e)
{
    for (auto &s : e->prof) {
        for (auto &ev : s.pending) {
            float ms = 0.0f;
            if (hipEventSynchronize(ev.second) == hipSuccess &&
                hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess) {
                s.total_ms += ms;
                s.launches += 1;
            }
            (void)hipEventDestroy(ev.first);
            (void)hipEventDestroy(ev.second);
        }
        s.pending.clear();
    }
}

bool drift_zero(const esrrf_engine *e, int c0, int n)
{
    const int N = e->p.n_frame_for_srrf;
    for (int i = c0; i < c0 + n; i++)
        if (e->h_drift[i] != 0.0f || e->h_drift[i + N] != 0.0f) return false;
    return true;
}

// RGC + accumulation for `n` frames of the current load starting at load index f0,
// SR-frame index c0 (BIC mode).
int run_bic_rgc(esrrf_engine *e, int f0, int c0, int n)
{
    const int N = e->p.n_frame_for_srrf;
    const float rn = 1.0f / (float)N;
    const float rn1 = 1.0f / (float)(N - 1);
    if (e->tiled && drift_zero(e, c0, n)) {
        BicRgcArgs a = bic_args(e->tabs, e->shape, e->d_U, e->d_px, e->w, e->h, e->M);
        a.out = e->d_out;
        a.prev = e->d_prev;
        a.frame0 = f0;
        a.nframes = n;
        a.first_init = (c0 == 0);
        a.rn = rn;
        a.rn1 = rn1;
        a.sens = e->p.sensitivity;
        a.sens_int = e->sens_int;
        a.int_weighting = e->p.int_weighting ? 1 : 0;
        a.stage = e->stage;
        a.pair_direct = e->pair_direct;
        ProfScope ps(e, 0);
        HIPCHK(e, launch_bic_rgc(a, e->shape.ppw, e->stream, &e->rgc_kernel));
        return ESRRF_OK;
    }
    GenericRgcArgs g{};
    g.U = e->d_U;
    g.px = e->d_px;
    g.drift = e->d_drift;
    g.out = e->d_out;
    g.prev = e->d_prev;
    g.w = e->w;
    g.h = e->h;
    g.M = e->M;
    g.wM = e->wM;
    g.hM = e->hM;
    g.frame0 = f0;
    g.nframes = n;
    g.count0 = c0;
    g.nFrameForSRRF = N;
    g.rn = rn;
    g.rn1 = rn1;
    g.sens = e->p.sensitivity;
    g.sens_int = e->sens_int;
    g.int_weighting = e->p.int_weighting ? 1 : 0;
    g.tac_prev_scaled = 1;
    g.radius = e->radius;
    g.tSS = e->tSS;
    g.tSO = e->tSO;
    ProfScope ps(e, 0);
    HIPCHK(e, launch_generic_rgc(g, e->stream));
    e->rgc_kernel = "k_generic_rgc";
    return ESRRF_OK;
}

// Smallest upload chunk worth its own RGC launch (frames).
constexpr int kMinUploadChunk = 64;
// Chunk sizes (even except the last, summing to n) of a load of n frames uploaded in at
// most nch chunks.  The kernels start once the first chunk is resident, so it is small; each
// chunk after it is three times the one before, which keeps its copy (about 2.6 us per 256^2
// uint16 frame over PCIe) shorter than the kernels of the previous chunk (about 10 us per
// frame at M = 5), so only the first copy is exposed.  The first chunk is the smallest that
// lets nch chunks cover the load, at least 16 frames.
std::vector<int> upload_schedule(int n, int nch)
{
    double tot = 0.0, g = 1.0;
    for (int k = 0; k < nch; k++, g *= 3.0) tot += g;
    long s = std::max(16L, (long)(n / tot));
    s += s & 1;
    std::vector<int> sz;
    int rem = n;
    for (int k = 0; rem > 0; k++) {
        int c = rem;
        if (k < nch - 1 && s < rem) c = (int)s;
        sz.push_back(c);
        rem -= c;
        s = std::min(3 * s, (long)n);
    }
    return sz;
}

// calculateSRRF for host frames in BIC mode with the upload overlapped: chunk c of the load
// is copied on copy_stream while the compute stream converts, preps and accumulates chunk
// c-1.  Each chunk's RGC launch continues the cumulants the previous one left in d_out /
// d_prev, as successive loads of one SR frame do, so the result equals one launch over the
// whole load bit for bit.  Chunks are issued copy-then-compute, so pageable host memory
// (whose copies return only when done) still overlaps the previous chunk's kernels.
int calculate_bic_pipelined(esrrf_engine *e, const void *frames, int dtype, int n)
{
    const size_t esz = dtype == ESRRF_U16 ? 2 : 4;
    const long npx = (long)e->W * e->H;
    const int nch = std::min(e->upload_chunks, n / kMinUploadChunk);
    // even chunks keep the frame-pair kernel on pairs
    const std::vector<int> sched = upload_schedule(n, nch);
    // the staging buffer may still be read by work queued earlier on the compute stream
    HIPCHK(e, hipEventRecord(e->ev_free, e->stream));
    HIPCHK(e, hipStreamWaitEvent(e->copy_stream, e->ev_free, 0));
    e->counters[1] = 0;
    e->loaded = n;
    int f0 = 0;
    for (int c = 0; c < (int)sched.size(); f0 += sched[c], c++) {
        const int k = sched[c];
        const char *hsrc = (const char *)frames + esz * npx * f0;
        char *dst = (char *)e->d_in + esz * npx * f0;
        HIPCHK(e, hipMemcpyAsync(dst, hsrc, esz * npx * k, hipMemcpyHostToDevice,
                                 e->copy_stream));
        HIPCHK(e, hipEventRecord(e->ev_chunk[c], e->copy_stream));
        HIPCHK(e, hipStreamWaitEvent(e->stream, e->ev_chunk[c], 0));
        HIPCHK(e, launch_convert_frames(dst, dtype, e->d_px + npx * f0, npx * k, e->stream));
        {
            ProfScope ps(e, 1);
            HIPCHK(e, launch_bic_prep(e->d_px + npx * f0, e->d_U + 4 * npx * f0, e->w, e->h, k,
                                      e->stream, e->var.bic_prep));
        }
        int rc = run_bic_rgc(e, f0, e->counters[0], k);
        if (rc != ESRRF_OK) return rc;
        e->counters[0] += k;
        e->counters[1] += k;
    }
    // the caller may reuse its frames on return, as after LiveSRRF_CL.java:711's finish;
    // the kernels keep running
    HIPCHK(e, hipStreamSynchronize(e->copy_stream));
    return ESRRF_OK;
}

int calculate_impl(esrrf_engine *e, const void *frames, bool on_device, int dtype, int n,
                   const volatile int *cancel, int *cancelled)
{
    if (cancelled) *cancelled = 0;
    if (!frames) return fail(e, ESRRF_ERR_INVALID, "frames is NULL");
    if (dtype != ESRRF_U16 && dtype != ESRRF_F32)
        return fail(e, ESRRF_ERR_INVALID, "dtype must be ESRRF_U16 or ESRRF_F32");
    if (n <= 0 || n > e->p.n_frames_on_gpu)
        return fail(e, ESRRF_ERR_INVALID, "n_frames=%d outside [1, n_frames_on_gpu=%d]", n,
                    e->p.n_frames_on_gpu);
    if (e->counters[0] + n > e->p.n_frame_for_srrf)
        return fail(e, ESRRF_ERR_STATE,
                    "frame %d of the SR frame exceeds n_frame_for_srrf=%d (call "
                    "esrrf_reset_frame_position between SR frames)",
                    e->counters[0] + n, e->p.n_frame_for_srrf);
    const size_t esz = dtype == ESRRF_U16 ? 2 : 4;
    const long npx = (long)e->W * e->H;
    const void *src = frames;
    if (!on_device && e->mode == MODE_BIC && !cancel && e->upload_chunks > 1 &&
        n >= 2 * kMinUploadChunk)
        return calculate_bic_pipelined(e, frames, dtype, n);
    if (!on_device) {
        HIPCHK(e, hipMemcpyAsync(e->d_in, frames, esz * npx * n, hipMemcpyHostToDevice,
                                 e->stream));
        HIPCHK(e, hipEventRecord(e->ev_chunk[0], e->stream));
        src = e->d_in;
    }
    e->counters[1] = 0;   // the gradient kernel resets the load counter (liveSRRF.cl:207)
    e->loaded = n;
    int rc = ESRRF_OK;
    if (e->mode == MODE_BIC) {
        HIPCHK(e, launch_convert_frames(src, dtype, e->d_px, npx * n, e->stream));
        {
            ProfScope ps(e, 1);
            HIPCHK(e, launch_bic_prep(e->d_px, e->d_U, e->w, e->h, n, e->stream,
                                      e->var.bic_prep));
        }
    } else {
        ModeCtx c = {e->stream, e->w, e->h, e->W, e->H, e->nP, e->M, e->wM, e->hM,
                     e->d_px, e->d_U, e->d_aux, e->profiling};
        rc = mode_prepare(e->mode == MODE_FHT ? 1 : 2, c, e->ms, src, dtype, n);
        if (rc != ESRRF_OK) return fail(e, rc, "%s", e->ms.err.c_str());
    }
    const int chunk = cancel ? 16 : n;
    for (int f0 = 0; f0 < n; f0 += chunk) {
        const int k = std::min(chunk, n - f0);
        if (e->mode == MODE_BIC) {
            rc = run_bic_rgc(e, f0, e->counters[0], k);
        } else {
            ModeCtx c = {e->stream, e->w, e->h, e->W, e->H, e->nP, e->M, e->wM, e->hM,
                         e->d_px, e->d_U, e->d_aux, e->profiling};
            ModeRgc r{};
            r.frame0 = f0;
            r.nframes = k;
            r.count0 = e->counters[0];
            r.nFrameForSRRF = e->p.n_frame_for_srrf;
            r.sens = e->p.sensitivity;
            r.sens_int = e->sens_int;
            r.int_weighting = e->p.int_weighting ? 1 : 0;
            r.radius = e->radius;
            r.tSS = e->tSS;
            r.tSO = e->tSO;
            r.drift = e->d_drift;
            r.drift_zero = drift_zero(e, e->counters[0], k);
            r.out = e->d_out;
            r.prev = e->d_prev;
            const int md = e->mode == MODE_FHT ? 1 : 2;
            rc = mode_pre_rgc(md, c, e->ms, r);
            if (rc == ESRRF_OK) {
                ProfScope ps(e, 0);
                rc = mode_rgc(md, c, e->ms, r);
            }
            if (rc != ESRRF_OK) rc = fail(e, rc, "%s", e->ms.err.c_str());
        }
        if (rc != ESRRF_OK) return rc;
        e->counters[0] += k;
        e->counters[1] += k;
        if (cancel && *cancel && f0 + k < n) {
            if (cancelled) *cancelled = 1;
            break;
        }
    }
    // the host frames are free for the caller on return (LiveSRRF_CL.java:711)
    if (!on_device) HIPCHK(e, hipEventSynchronize(e->ev_chunk[0]));
    return ESRRF_OK;
}

// VAR, then the MP map and correction (channels 0..2; LiveSRRF_CL.java:757-770)
int epilogue_impl(esrrf_engine *e)
{
    HIPCHK(e, launch_var(e->d_out, e->nvox, e->stream));
    if (e->p.do_mp_map_correction && e->mode != MODE_FHT) {
        HIPCHK(e, launch_mp_map(e->d_out, e->d_mp, e->w, e->h, e->nP, e->M, e->Mz, e->stream));
        HIPCHK(e, launch_mp_correct(e->d_out, e->d_mp, e->w, e->h, e->nP, e->M, e->Mz, e->stream));
        e->has_mp = true;
    }
    return ESRRF_OK;
}

int finalize_impl(esrrf_engine *e, float *dst_dev)
{
    const int rc = epilogue_impl(e);
    if (rc) return rc;
    HIPCHK(e, launch_nan_to_zero_copy(e->d_out, dst_dev, 4 * e->nvox, e->stream));
    return ESRRF_OK;
}

}  // namespace

extern "C" {

int esrrf_device_count(int *n)
{
    if (!n) return ESRRF_ERR_INVALID;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return ESRRF_OK;
}

int esrrf_device_name(int device, char *buf, int len)
{
    if (!buf || len <= 0) return ESRRF_ERR_INVALID;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return ESRRF_ERR_INVALID;
    snprintf(buf, (size_t)len, "%s (%s)", prop.name, prop.gcnArchName);
    return ESRRF_OK;
}

int esrrf_fastest_device(int *device)
{
    if (!device) return ESRRF_ERR_INVALID;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return ESRRF_ERR_HIP;
    double best = -1;
    int bi = 0;
    for (int i = 0; i < n; i++) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, i) != hipSuccess) continue;
        double fl = (double)prop.multiProcessorCount * prop.clockRate;
        if (fl > best) {
            best = fl;
            bi = i;
        }
    }
    *device = bi;
    return ESRRF_OK;
}

int esrrf_create(const esrrf_params *p, esrrf_engine **out)
{
    if (!p || !out) return ESRRF_ERR_INVALID;
    *out = nullptr;
    esrrf_engine *e = new esrrf_engine();
    e->p = *p;
    auto bad = [&](int code) {
        (void)hipGetLastError();   // a failed call must not surface in later checks
        free_all(e);
        std::string msg = e->err;
        delete e;
        (void)msg;
        return code;
    };
    if (p->width <= 0 || p->height <= 0 || p->magnification < 1 || p->n_frames_on_gpu < 1 ||
        p->n_frame_for_srrf < 1 || !(p->fwhm > 0.0f) || !(p->sensitivity == p->sensitivity) ||
        p->axial_magnification < 0) {
        return bad(fail(e, ESRRF_ERR_INVALID, "invalid parameters"));
    }
    int dev = p->device;
    if (dev < 0 && esrrf_fastest_device(&dev) != ESRRF_OK) return bad(ESRRF_ERR_HIP);
    e->dev = dev;
    if (hipSetDevice(dev) != hipSuccess) return bad(ESRRF_ERR_HIP);
    if (hipStreamCreate(&e->own_stream) != hipSuccess) return bad(ESRRF_ERR_HIP);
    e->stream = e->own_stream;
    if (hipStreamCreate(&e->copy_stream) != hipSuccess) return bad(ESRRF_ERR_HIP);
    if (hipEventCreateWithFlags(&e->ev_free, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_int, hipEventDisableTiming) != hipSuccess)
        return bad(ESRRF_ERR_HIP);
    e->ev_chunk.assign(kMaxUploadChunks, nullptr);
    for (hipEvent_t &ev : e->ev_chunk)
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
            return bad(ESRRF_ERR_HIP);
    const bool do3d = p->n_planes > 0;
    e->mode = do3d ? MODE_3D : (p->do_fht_interpolation ? MODE_FHT : MODE_BIC);
    e->M = p->magnification;
    e->W = p->width;
    e->H = p->height;
    if (do3d) {
        if (!p->shift_x_3d || !p->shift_y_3d || !p->theta_3d || !p->roi_3d || !p->int_coeffs_3d)
            return bad(ESRRF_ERR_INVALID);
        e->nP = p->n_planes;
        e->w = p->width / 3;    // nSplits = 3 (LiveSRRF_CL.java:39,229-230)
        e->h = p->height / 3;
        e->Mz = p->axial_magnification > 0 ? p->axial_magnification : e->M;
        e->nPlanesM = e->nP * e->Mz;
    } else {
        e->nP = 1;
        e->w = p->width;
        e->h = p->height;
        e->nPlanesM = 1;
    }
    if (e->w < 4 || e->h < 4)
        return bad(fail(e, ESRRF_ERR_INVALID, "frames must be at least 4x4 pixels (per plane)"));
    e->wM = e->w * e->M;
    e->hM = e->h * e->M;
    e->nvox = (long)e->wM * e->hM * e->nPlanesM;
    e->sens_int = sens_exact_int(p->sensitivity);
    e->var = kernel_variants_from_env();
    if (e->mode == MODE_FHT)
        esrrf_fht_constants(p->fwhm, e->M, &e->sigma, &e->radius, &e->tSS, &e->tSO);
    else
        esrrf_bic_constants(p->fwhm, &e->sigma, &e->radius, &e->tSS, &e->tSO);

    const int nF = p->n_frames_on_gpu;
    const long npx = (long)e->W * e->H;
    e->d_in_bytes = (size_t)npx * nF * 4;
    int rc;
    if ((rc = dalloc(e, &e->d_in, e->d_in_bytes))) return bad(rc);
    if ((rc = dalloc(e, &e->d_out, sizeof(float) * 4 * e->nvox))) return bad(rc);
    if ((rc = dalloc(e, &e->d_tmp, sizeof(float) * 4 * e->nvox))) return bad(rc);
    if ((rc = dalloc(e, &e->d_prev, sizeof(float) * e->nvox))) return bad(rc);
    if ((rc = dalloc(e, &e->d_drift, sizeof(float) * 2 * p->n_frame_for_srrf))) return bad(rc);
    const int Md = e->M * e->M * e->Mz;
    if ((rc = dalloc(e, &e->d_mp, sizeof(float) * 3 * Md))) return bad(rc);
    // defined initial state (the reference leaves fresh OpenCL buffers uninitialised)
    if (hipMemsetAsync(e->d_out, 0, sizeof(float) * 4 * e->nvox, e->stream) != hipSuccess ||
        hipMemsetAsync(e->d_prev, 0, sizeof(float) * e->nvox, e->stream) != hipSuccess ||
        hipMemsetAsync(e->d_drift, 0, sizeof(float) * 2 * p->n_frame_for_srrf, e->stream) !=
            hipSuccess ||
        hipMemsetAsync(e->d_mp, 0, sizeof(float) * 3 * Md, e->stream) != hipSuccess)
        return bad(fail(e, ESRRF_ERR_HIP, "hipMemsetAsync failed"));
    e->h_drift.assign(2 * (size_t)p->n_frame_for_srrf, 0.0f);

    if (e->mode == MODE_BIC) {
        if ((rc = dalloc(e, &e->d_px, sizeof(float) * npx * nF))) return bad(rc);
        if ((rc = dalloc(e, &e->d_U, sizeof(float2) * 4 * npx * nF))) return bad(rc);
        if (bic_tables_create(e->w, e->h, e->M, p->fwhm, e->tabs) != hipSuccess)
            return bad(fail(e, ESRRF_ERR_HIP, "tap table upload failed"));
        e->tiled = e->tabs.ok;
        if (e->tiled) {
            e->shape = bic_default_shape(e->tabs, e->M);
            e->stage = e->var.bic_stage;
            e->pair_direct = e->var.bic_pair_direct;
        }
    } else {
        ModeCreate mc{};
        mc.mode = e->mode == MODE_FHT ? 1 : 2;
        mc.w = e->w;
        mc.h = e->h;
        mc.W = e->W;
        mc.H = e->H;
        mc.nP = e->nP;
        mc.M = e->M;
        mc.Mz = e->Mz;
        mc.nF = nF;
        mc.radius = e->radius;
        mc.tSS = e->tSS;
        mc.tSO = e->tSO;
        mc.p = p;
        mc.var = e->var;
        rc = mode_create(mc, e->ms, &e->d_px, &e->d_U, &e->d_aux, &e->d_aux_bytes);
        if (rc != ESRRF_OK) {
            fail(e, rc, "%s", e->ms.err.c_str());
            return bad(rc);
        }
    }
    if (hipStreamSynchronize(e->stream) != hipSuccess) return bad(ESRRF_ERR_HIP);
    *out = e;
    return ESRRF_OK;
}

int esrrf_destroy(esrrf_engine *e)
{
    if (!e) return ESRRF_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        (void)hipSetDevice(e->dev);
        (void)hipStreamSynchronize(e->stream);
        free_all(e);
    }
    delete e;
    return ESRRF_OK;
}

int esrrf_load_drift(esrrf_engine *e, const float *dx, const float *dy, int n)
{
    if (!e || !dx || !dy) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    const int N = e->p.n_frame_for_srrf;
    if (n < N) return fail(e, ESRRF_ERR_INVALID, "drift arrays hold %d < n_frame_for_srrf=%d", n, N);
    // driftXY[i] = magnification * driftX[i] (LiveSRRF_CL.java:586-589), in all modes
    for (int i = 0; i < N; i++) {
        e->h_drift[i] = (float)e->M * dx[i];
        e->h_drift[i + N] = (float)e->M * dy[i];
    }
    (void)hipSetDevice(e->dev);
    HIPCHK(e, hipMemcpyAsync(e->d_drift, e->h_drift.data(), sizeof(float) * 2 * N,
                             hipMemcpyHostToDevice, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return ESRRF_OK;
}

int esrrf_reset_frame_position(esrrf_engine *e)
{
    if (!e) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    e->counters[0] = 0;
    return ESRRF_OK;
}

int esrrf_calculate(esrrf_engine *e, const void *frames, int dtype, int n_frames,
                    const volatile int *cancel, int *cancelled)
{
    if (!e) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    (void)hipSetDevice(e->dev);
    return calculate_impl(e, frames, false, dtype, n_frames, cancel, cancelled);
}

int esrrf_calculate_device(esrrf_engine *e, const void *d_frames, int dtype, int n_frames,
                           const volatile int *cancel, int *cancelled)
{
    if (!e) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    (void)hipSetDevice(e->dev);
    return calculate_impl(e, d_frames, true, dtype, n_frames, cancel, cancelled);
}

int esrrf_read_result(esrrf_engine *e, float *out)
{
    if (!e || !out) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    (void)hipSetDevice(e->dev);
    // INT (channel 3) is final once the RGC launches are done: VAR and the MP correction
    // change channels 0..2 only (LiveSRRF_CL.java:757-770).  Its NaN-to-0 pass runs first and
    // its copy to the host goes on the copy stream while the epilogue kernels finish the other
    // three channels; the values are those of one pass over all four.
    const long nv = e->nvox;
    HIPCHK(e, launch_nan_to_zero_copy(e->d_out + 3 * nv, e->d_tmp + 3 * nv, nv, e->stream));
    HIPCHK(e, hipEventRecord(e->ev_int, e->stream));
    const int rc = epilogue_impl(e);
    if (rc) return rc;
    HIPCHK(e, launch_nan_to_zero_copy(e->d_out, e->d_tmp, 3 * nv, e->stream));
    // (the epilogue is queued before the first copy: a copy to pageable memory can return
    // only once it is done; both streams are drained on every path, so no copy into `out`
    // outlives the call)
    hipError_t st = hipStreamWaitEvent(e->copy_stream, e->ev_int, 0);
    if (st == hipSuccess)
        st = hipMemcpyAsync(out + 3 * nv, e->d_tmp + 3 * nv, sizeof(float) * nv,
                            hipMemcpyDeviceToHost, e->copy_stream);
    if (st == hipSuccess)
        st = hipMemcpyAsync(out, e->d_tmp, sizeof(float) * 3 * nv, hipMemcpyDeviceToHost,
                            e->stream);
    const hipError_t s1 = hipStreamSynchronize(e->copy_stream);
    const hipError_t s2 = hipStreamSynchronize(e->stream);
    if (st == hipSuccess) st = s1 != hipSuccess ? s1 : s2;
    if (st != hipSuccess)
        return fail(e, ESRRF_ERR_HIP, "result read-back: %s", hipGetErrorString(st));
    return ESRRF_OK;
}

int esrrf_read_result_device(esrrf_engine *e, float *d_out)
{
    if (!e || !d_out) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    (void)hipSetDevice(e->dev);
    return finalize_impl(e, d_out);
}

int esrrf_read_gradients(esrrf_engine *e, float *out)
{
    if (!e || !out) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    (void)hipSetDevice(e->dev);
    if (e->loaded == 0) return fail(e, ESRRF_ERR_STATE, "no frames loaded");
    if (e->mode == MODE_BIC) {
        const long n = 4L * e->w * e->h;
        float *tmp = nullptr;
        int rc = dalloc(e, &tmp, sizeof(float) * 2 * n);
        if (rc) return rc;
        hipError_t st = launch_bic_gint(e->d_px, tmp, tmp + n, e->w, e->h, 0, e->stream);
        if (st == hipSuccess) st = launch_nan_to_zero_copy(tmp, tmp, 2 * n, e->stream);
        if (st == hipSuccess)
            st = hipMemcpyAsync(out, tmp, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, e->stream);
        if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
        (void)hipFree(tmp);
        if (st != hipSuccess) return fail(e, ESRRF_ERR_HIP, "gradient read-back: %s", hipGetErrorString(st));
        return ESRRF_OK;
    }
    ModeCtx c = {e->stream, e->w, e->h, e->W, e->H, e->nP, e->M, e->wM, e->hM,
                 e->d_px, e->d_U, e->d_aux, e->profiling};
    int rc = mode_read_gradients(e->mode == MODE_FHT ? 1 : 2, c, e->ms, out);
    if (rc) return fail(e, rc, "%s", e->ms.err.c_str());
    return ESRRF_OK;
}

int esrrf_prepare_fht(esrrf_engine *e, const void *frames, int dtype, int n_frames)
{
    if (!e || !frames) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    (void)hipSetDevice(e->dev);
    if (e->mode != MODE_FHT)
        return fail(e, ESRRF_ERR_STATE, "prepareDataFHT_SRRF exists only in FHT mode");
    if (dtype != ESRRF_U16 && dtype != ESRRF_F32)
        return fail(e, ESRRF_ERR_INVALID, "dtype must be ESRRF_U16 or ESRRF_F32");
    if (n_frames <= 0 || n_frames > e->p.n_frames_on_gpu)
        return fail(e, ESRRF_ERR_INVALID, "n_frames=%d outside [1, n_frames_on_gpu=%d]",
                    n_frames, e->p.n_frames_on_gpu);
    const size_t esz = dtype == ESRRF_U16 ? 2 : 4;
    HIPCHK(e, hipMemcpyAsync(e->d_in, frames, esz * (size_t)e->W * e->H * n_frames,
                             hipMemcpyHostToDevice, e->stream));
    ModeCtx c = {e->stream, e->w, e->h, e->W, e->H, e->nP, e->M, e->wM, e->hM,
                 e->d_px, e->d_U, e->d_aux, e->profiling};
    const int rc = mode_prepare(1, c, e->ms, e->d_in, dtype, n_frames);
    if (rc != ESRRF_OK) return fail(e, rc, "%s", e->ms.err.c_str());
    e->counters[1] = 0;   // the gradient kernel resets the load counter (liveSRRF.cl:207)
    e->loaded = n_frames;
    HIPCHK(e, hipStreamSynchronize(e->stream));   // the host frames are free on return
    return ESRRF_OK;
}

int esrrf_reshape_3d(esrrf_engine *e, const void *frames, int dtype, int n_frames, float *out)
{
    if (!e || !frames || !out) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    (void)hipSetDevice(e->dev);
    if (e->mode != MODE_3D)
        return fail(e, ESRRF_ERR_STATE, "reshapeImageStack3D exists only in 3-D mode");
    if (dtype != ESRRF_U16 && dtype != ESRRF_F32)
        return fail(e, ESRRF_ERR_INVALID, "dtype must be ESRRF_U16 or ESRRF_F32");
    if (n_frames <= 0 || n_frames > e->p.n_frames_on_gpu)
        return fail(e, ESRRF_ERR_INVALID, "n_frames=%d outside [1, n_frames_on_gpu=%d]",
                    n_frames, e->p.n_frames_on_gpu);
    const size_t esz = dtype == ESRRF_U16 ? 2 : 4;
    HIPCHK(e, hipMemcpyAsync(e->d_in, frames, esz * (size_t)e->W * e->H * n_frames,
                             hipMemcpyHostToDevice, e->stream));
    ModeCtx c = {e->stream, e->w, e->h, e->W, e->H, e->nP, e->M, e->wM, e->hM,
                 e->d_px, e->d_U, e->d_aux, e->profiling};
    const int rc = mode_reshape_3d(c, e->ms, e->d_in, dtype, n_frames, out);
    if (rc != ESRRF_OK) return fail(e, rc, "%s", e->ms.err.c_str());
    return ESRRF_OK;
}

int esrrf_read_aligned(esrrf_engine *e, float *out)
{
    if (!e || !out) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->mode != MODE_3D) return fail(e, ESRRF_ERR_STATE, "aligned pixels exist only in 3-D mode");
    if (e->loaded == 0) return fail(e, ESRRF_ERR_STATE, "no frames loaded");
    (void)hipSetDevice(e->dev);
    const long n = (long)e->w * e->h * e->nP;
    HIPCHK(e, hipMemcpyAsync(out, e->d_px, sizeof(float) * n, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    for (long i = 0; i < n; i++)
        if (out[i] != out[i]) out[i] = 0.0f;
    return ESRRF_OK;
}

int esrrf_read_mp_maps(esrrf_engine *e, float *out)
{
    if (!e || !out) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->mode == MODE_FHT)
        return fail(e, ESRRF_ERR_STATE, "no MP map in FHT mode (LiveSRRF_CL.java:762)");
    (void)hipSetDevice(e->dev);
    const int Md = e->M * e->M * e->Mz;
    HIPCHK(e, hipMemcpyAsync(out, e->d_mp, sizeof(float) * 3 * Md, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(e, hipStreamSynchronize(e->stream));
    for (int i = 0; i < 3 * Md; i++)
        if (out[i] != out[i]) out[i] = 0.0f;
    return ESRRF_OK;
}

int esrrf_read_frame_counters(esrrf_engine *e, int *out2)
{
    if (!e || !out2) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    out2[0] = e->counters[0];
    out2[1] = e->counters[1];
    return ESRRF_OK;
}

int esrrf_get_dims(esrrf_engine *e, int *width_m, int *height_m, int *n_planes_m, int *width_s,
                   int *height_s, int *n_planes)
{
    if (!e) return ESRRF_ERR_INVALID;
    if (width_m) *width_m = e->wM;
    if (height_m) *height_m = e->hM;
    if (n_planes_m) *n_planes_m = e->nPlanesM;
    if (width_s) *width_s = e->w;
    if (height_s) *height_s = e->h;
    if (n_planes) *n_planes = e->nP;
    return ESRRF_OK;
}

int esrrf_finish(esrrf_engine *e)
{
    if (!e) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    (void)hipSetDevice(e->dev);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    return ESRRF_OK;
}

int esrrf_set_stream(esrrf_engine *e, void *hip_stream)
{
    if (!e) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    (void)hipSetDevice(e->dev);
    HIPCHK(e, hipStreamSynchronize(e->stream));
    e->stream = hip_stream ? (hipStream_t)hip_stream : e->own_stream;
    return ESRRF_OK;
}

int esrrf_host_alloc(size_t bytes, void **ptr)
{
    if (!ptr || bytes == 0) return ESRRF_ERR_INVALID;
    *ptr = nullptr;
    if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess) {
        *ptr = nullptr;
        (void)hipGetLastError();
        return ESRRF_ERR_NOMEM;
    }
    return ESRRF_OK;
}

int esrrf_host_free(void *ptr)
{
    if (!ptr) return ESRRF_OK;
    return hipHostFree(ptr) == hipSuccess ? ESRRF_OK : ESRRF_ERR_HIP;
}

int esrrf_set_upload_chunks(esrrf_engine *e, int chunks)
{
    if (!e || chunks < 1 || chunks > kMaxUploadChunks) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    e->upload_chunks = chunks;
    return ESRRF_OK;
}

int esrrf_profile_enable(esrrf_engine *e, int enable)
{
    if (!e) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    prof_collect(e);
    for (auto &s : e->prof) {
        s.total_ms = 0.0;
        s.launches = 0;
    }
    e->profiling = enable != 0;
    return ESRRF_OK;
}

int esrrf_profile_read(esrrf_engine *e, int which, double *total_ms, long *launches)
{
    if (!e || which < 0 || which > 2) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    (void)hipSetDevice(e->dev);
    prof_collect(e);
    ProfSlot &s = e->prof[which];
    if (which == 2) {
        s.total_ms = e->ms.prof_up_ms;
        s.launches = e->ms.prof_up_n;
    }
    if (total_ms) *total_ms = s.total_ms;
    if (launches) *launches = s.launches;
    return ESRRF_OK;
}

int esrrf_active_samples(esrrf_engine *e, long long *per_frame)
{
    if (!e || !per_frame) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->mode == MODE_BIC)
        *per_frame = count_active_samples_bic(e->w, e->h, e->M, e->radius, e->tSO);
    else
        *per_frame = mode_active_samples(e->mode == MODE_FHT ? 1 : 2, e->ms);
    return ESRRF_OK;
}

int esrrf_autotune(esrrf_engine *e, int reps, int *nw, int *ppw, int *taps_in_lds, float *ms)
{
    if (!e) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->mode != MODE_BIC || !e->tiled)
        return fail(e, ESRRF_ERR_STATE, "autotuning applies to the tiled 2-D BIC kernel only");
    if (e->loaded == 0) return fail(e, ESRRF_ERR_STATE, "no frames loaded");
    (void)hipSetDevice(e->dev);
    // candidates run over the whole current load into scratch outputs: the engine's
    // cumulants and frame counters are left untouched
    float *prev = nullptr;
    int rc = dalloc(e, &prev, sizeof(float) * e->nvox);
    if (rc) return rc;
    BicRgcArgs a = bic_args(e->tabs, e->shape, e->d_U, e->d_px, e->w, e->h, e->M);
    a.out = e->d_tmp;
    a.prev = prev;
    a.frame0 = 0;
    a.nframes = e->loaded;
    a.first_init = 1;
    a.rn = 1.0f / (float)e->p.n_frame_for_srrf;
    a.rn1 = 1.0f / (float)(e->p.n_frame_for_srrf - 1);
    a.sens = e->p.sensitivity;
    a.sens_int = e->sens_int;
    a.int_weighting = e->p.int_weighting ? 1 : 0;
    BicShape best;
    hipError_t st = bic_autotune(e->tabs, a, e->M, std::max(reps, 1), e->stream, &best, nullptr);
    (void)hipFree(prev);
    if (st != hipSuccess) return fail(e, ESRRF_ERR_HIP, "autotune: %s", hipGetErrorString(st));
    e->shape = best;
    if (nw) *nw = best.nw;
    if (ppw) *ppw = best.ppw;
    if (taps_in_lds) *taps_in_lds = best.taps_in_lds;
    if (ms) *ms = best.ms;
    return ESRRF_OK;
}

int esrrf_set_launch_shape(esrrf_engine *e, int nw, int taps_in_lds, int frames_per_pass)
{
    if (!e) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->mode != MODE_BIC || !e->tiled)
        return fail(e, ESRRF_ERR_STATE, "no tiled 2-D kernel in this mode or geometry");
    BicShape s;
    s.nw = nw;
    s.ppw = nw >= 1 ? (e->M * e->M + nw - 1) / nw : 0;
    s.taps_in_lds = taps_in_lds ? 1 : 0;
    s.fp = frames_per_pass;
    if (!bic_shape_supported(e->tabs, e->M, s))
        return fail(e, ESRRF_ERR_UNSUPPORTED,
                    "launch shape nw=%d taps_in_lds=%d frames_per_pass=%d not runnable here", nw,
                    taps_in_lds, frames_per_pass);
    e->shape = s;
    return ESRRF_OK;
}

int esrrf_launch_shape(esrrf_engine *e, int *nw, int *ppw, int *taps_in_lds,
                       int *frames_per_pass)
{
    if (!e) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->mode != MODE_BIC || !e->tiled)
        return fail(e, ESRRF_ERR_STATE, "the engine runs no tiled 2-D kernel");
    if (nw) *nw = e->shape.nw;
    if (ppw) *ppw = e->shape.ppw;
    if (taps_in_lds) *taps_in_lds = e->shape.taps_in_lds;
    if (frames_per_pass) *frames_per_pass = e->shape.fp;
    return ESRRF_OK;
}

int esrrf_rgc_kernel(esrrf_engine *e, char *buf, int len)
{
    if (!e || !buf || len < 1) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(e->mu);
    const char *k = e->mode == MODE_BIC ? e->rgc_kernel : e->ms.rgc_kernel;
    snprintf(buf, (size_t)len, "%s", k);
    return ESRRF_OK;
}

const char *esrrf_last_error(esrrf_engine *e)
{
    if (!e) return "null engine";
    return e->err.c_str();
}

const char *esrrf_version(void)
{
    return "esrrf-mi355x 0.1 (gfx950)";
}

}  // extern "C"
