// drift.cpp -- drift of a window's frames against its first frame
// (XYShiftCalculator.calculateShiftArray, XYShiftCalculator.java:25-47), for
// loadDriftXYGPUbuffer.  The reference reads the shift from NanoJ-Core's cross-correlation
// map and peak fit, which are not part of the reference tree; this is the standard method:
// mean-free frames zero-padded to the next power-of-two square, circular cross-correlation
// by FFT, the first maximum and a three-point parabola per axis.  The result is the
// displacement of each frame's content relative to the reference (f(x) = r(x - d) gives d).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/esrrf.h"
#include "drift_internal.h"
#include "frc_internal.h"

using namespace esrrf;

struct esrrf_drift {
    std::mutex mu;
    std::string err;
    int dev = 0;
    hipStream_t own_stream = nullptr, stream = nullptr;
    DriftPlan plan{};
    float *d_shift = nullptr;    // [2][max_batch]
    float *d_stage = nullptr;    // host-call staging of the frames
    size_t stage_frames = 0;
    std::vector<float> h_shift;
};

namespace {

int dfail(esrrf_drift *q, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (q) q->err = buf;
    return code;
}

// a failed call also leaves the runtime's last error set: cleared, so that later launch
// checks on this thread do not report it as theirs
#define DCHK(q, x)                                                                         \
    do {                                                                                   \
        hipError_t _st = (x);                                                              \
        if (_st != hipSuccess) {                                                           \
            (void)hipGetLastError();                                                       \
            return dfail((q), ESRRF_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(_st)); \
        }                                                                                  \
    } while (0)

void drift_free(esrrf_drift *q)
{
    DriftPlan &p = q->plan;
    void *ptrs[] = {(void *)p.tw, p.R, p.buf, p.sums, p.pkv, p.pki, q->d_shift, q->d_stage};
    for (void *x : ptrs)
        if (x) (void)hipFree(x);
    if (q->own_stream) (void)hipStreamDestroy(q->own_stream);
}

int run_impl(esrrf_drift *q, const float *d_frames, int n, int ref, float *sx, float *sy)
{
    const DriftPlan &p = q->plan;
    const long wh = (long)p.w * p.h;
    DCHK(q, launch_drift_reference(p, d_frames + wh * ref, q->stream));
    for (int b0 = 0; b0 < n; b0 += p.max_batch) {
        const int nb = std::min(p.max_batch, n - b0);
        DCHK(q, launch_drift_batch(p, d_frames + wh * b0, nb, q->d_shift,
                                   q->d_shift + p.max_batch, q->stream));
        DCHK(q, hipMemcpyAsync(q->h_shift.data(), q->d_shift, sizeof(float) * 2 * p.max_batch,
                               hipMemcpyDeviceToHost, q->stream));
        DCHK(q, hipStreamSynchronize(q->stream));
        for (int k = 0; k < nb; k++) {
            if (sx) sx[b0 + k] = q->h_shift[k];
            if (sy) sy[b0 + k] = q->h_shift[p.max_batch + k];
        }
    }
    return ESRRF_OK;
}

}  // namespace

int esrrf_drift_create(int width, int height, int device, esrrf_drift **out)
{
    if (!out || width < 4 || height < 4) return ESRRF_ERR_INVALID;
    *out = nullptr;
    int N = 4, lg = 2;
    while (N < std::max(width, height)) {
        N <<= 1;
        lg++;
    }
    if (N > FRC_MAX_N) return ESRRF_ERR_UNSUPPORTED;
    esrrf_drift *q = new esrrf_drift();
    auto bad = [&](int code) {
        (void)hipGetLastError();   // a failed call must not surface in later checks
        drift_free(q);
        delete q;
        return code;
    };
    int dev = device;
    if (dev < 0 && esrrf_fastest_device(&dev) != ESRRF_OK) return bad(ESRRF_ERR_HIP);
    q->dev = dev;
    if (hipSetDevice(dev) != hipSuccess) return bad(ESRRF_ERR_HIP);
    if (hipStreamCreate(&q->own_stream) != hipSuccess) return bad(ESRRF_ERR_HIP);
    q->stream = q->own_stream;
    DriftPlan &p = q->plan;
    p.w = width;
    p.h = height;
    p.N = N;
    p.log2N = lg;
    // transforms of up to 64 frames or 64 MiB per batch
    p.max_batch = (int)std::max<long>(1, std::min<long>(64, (64L << 20) / (8L * N * N)));
    std::vector<float2> tw(N);
    for (int k = 0; k < N; k++) {
        const double ang = -2.0 * M_PI * k / N;
        tw[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    float2 *d_tw = nullptr;
    if (hipMalloc((void **)&d_tw, sizeof(float2) * N) != hipSuccess) return bad(ESRRF_ERR_NOMEM);
    p.tw = d_tw;
    if (hipMalloc((void **)&p.R, sizeof(float2) * (size_t)N * N) != hipSuccess ||
        hipMalloc((void **)&p.buf, sizeof(float2) * (size_t)p.max_batch * N * N) != hipSuccess ||
        hipMalloc((void **)&p.sums, sizeof(double) * (size_t)p.max_batch * (1 + DRIFT_SEG)) !=
            hipSuccess ||
        hipMalloc((void **)&p.pkv, sizeof(float) * (size_t)p.max_batch * DRIFT_SEG) != hipSuccess ||
        hipMalloc((void **)&p.pki, sizeof(int) * (size_t)p.max_batch * DRIFT_SEG) != hipSuccess ||
        hipMalloc((void **)&q->d_shift, sizeof(float) * 2 * (size_t)p.max_batch) != hipSuccess)
        return bad(ESRRF_ERR_NOMEM);
    if (hipMemcpy(d_tw, tw.data(), sizeof(float2) * N, hipMemcpyHostToDevice) != hipSuccess)
        return bad(ESRRF_ERR_HIP);
    q->h_shift.resize(2 * (size_t)p.max_batch);
    *out = q;
    return ESRRF_OK;
}

int esrrf_drift_destroy(esrrf_drift *q)
{
    if (!q) return ESRRF_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(q->mu);
        (void)hipSetDevice(q->dev);
        (void)hipStreamSynchronize(q->stream);
        drift_free(q);
    }
    delete q;
    return ESRRF_OK;
}

int esrrf_drift_set_stream(esrrf_drift *q, void *hip_stream)
{
    if (!q) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(q->mu);
    q->stream = hip_stream ? (hipStream_t)hip_stream : q->own_stream;
    return ESRRF_OK;
}

int esrrf_drift_run_device(esrrf_drift *q, const float *d_frames, int n_frames, int ref_index,
                           float *shift_x, float *shift_y)
{
    if (!q) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(q->mu);
    if (!d_frames || n_frames < 1 || ref_index < 0 || ref_index >= n_frames)
        return dfail(q, ESRRF_ERR_INVALID, "frames, n_frames >= 1 and 0 <= ref_index < n_frames");
    DCHK(q, hipSetDevice(q->dev));
    return run_impl(q, d_frames, n_frames, ref_index, shift_x, shift_y);
}

int esrrf_drift_run(esrrf_drift *q, const float *frames, int n_frames, int ref_index,
                    float *shift_x, float *shift_y)
{
    if (!q) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(q->mu);
    if (!frames || n_frames < 1 || ref_index < 0 || ref_index >= n_frames)
        return dfail(q, ESRRF_ERR_INVALID, "frames, n_frames >= 1 and 0 <= ref_index < n_frames");
    DCHK(q, hipSetDevice(q->dev));
    const size_t wh = (size_t)q->plan.w * q->plan.h;
    if (q->stage_frames < (size_t)n_frames) {
        if (q->d_stage) DCHK(q, hipFree(q->d_stage));
        q->d_stage = nullptr;
        q->stage_frames = 0;
        DCHK(q, hipMalloc((void **)&q->d_stage, sizeof(float) * wh * n_frames));
        q->stage_frames = (size_t)n_frames;
    }
    DCHK(q, hipMemcpyAsync(q->d_stage, frames, sizeof(float) * wh * n_frames,
                           hipMemcpyHostToDevice, q->stream));
    return run_impl(q, q->d_stage, n_frames, ref_index, shift_x, shift_y);
}

const char *esrrf_drift_last_error(esrrf_drift *q) { return q ? q->err.c_str() : ""; }
