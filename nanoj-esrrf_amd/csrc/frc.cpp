// frc.cpp -- Fourier ring correlation resolution of two images (the sweep's
// FRC.calculateFireNumber(odd, even, FIXED_1_OVER_7), gui/ParametersSweep_.java:437-446).
// NanoJ-Core's FRC is not part of the reference tree; this is the standard definition:
// centred square crop, Tukey(0.25) taper, zero padding to a power of two, ring r = the
// nearest integer to |k|, FRC(r) = sum Re(F1 conj F2) / sqrt(sum|F1|^2 sum|F2|^2), and the
// resolution N / r* at the first crossing r* of 1/7 after ring 0 (linear between rings).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/esrrf.h"
#include "frc_internal.h"

using namespace esrrf;

struct esrrf_frc {
    std::mutex mu;
    std::string err;
    int dev = 0;
    hipStream_t own_stream = nullptr, stream = nullptr;
    FrcPlan plan{};
    float *d_stage = nullptr;   // host-call staging [2][H][W]
    std::vector<double> h_rings;
};

namespace {

int ffail(esrrf_frc *q, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (q) q->err = buf;
    return code;
}

#define FCHK(q, x)                                                                         \
    do {                                                                                   \
        hipError_t _st = (x);                                                              \
        if (_st != hipSuccess)                                                             \
            return ffail((q), ESRRF_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(_st)); \
    } while (0)

void frc_free(esrrf_frc *q)
{
    FrcPlan &p = q->plan;
    void *ptrs[] = {(void *)p.win, (void *)p.tw, p.buf, p.part, p.rings, q->d_stage};
    for (void *x : ptrs)
        if (x) (void)hipFree(x);
    if (q->own_stream) (void)hipStreamDestroy(q->own_stream);
}

// the Tukey window of n points, alpha 0.25, computed in double
std::vector<float> tukey(int n)
{
    const double alpha = 0.25;
    std::vector<float> w(n, 1.0f);
    if (n == 1) return w;
    for (int i = 0; i < n; i++) {
        const double x = (double)i / (n - 1);
        double v = 1.0;
        if (x < alpha / 2) v = 0.5 * (1.0 + std::cos(2.0 * M_PI / alpha * (x - alpha / 2)));
        else if (x > 1.0 - alpha / 2)
            v = 0.5 * (1.0 + std::cos(2.0 * M_PI / alpha * (x - 1.0 + alpha / 2)));
        w[i] = (float)v;
    }
    return w;
}

}  // namespace

double esrrf::frc_resolution(const double *rings, int N, float *curve)
{
    const int R = N / 2;
    const double *num = rings, *d1 = rings + R, *d2 = rings + 2 * R;
    const double t = 1.0 / 7.0;
    double res = std::numeric_limits<double>::quiet_NaN(), prev = 0.0;
    bool found = false;
    for (int r = 0; r < R; r++) {
        const double den = std::sqrt(d1[r] * d2[r]);
        const double c = den > 0.0 ? num[r] / den : 0.0;
        if (curve) curve[r] = (float)c;
        if (!found && r >= 1 && c < t) {
            const double rs = prev >= t ? (r - 1) + (prev - t) / (prev - c) : r;
            res = (double)N / rs;
            found = true;
        }
        prev = c;
    }
    return res;
}

namespace {

int run_impl(esrrf_frc *q, const float *d_a, const float *d_b, double *resolution, float *curve,
             int *n_rings)
{
    FrcPlan &p = q->plan;
    FCHK(q, launch_frc(p, d_a, d_b, q->stream));
    const int R = p.N / 2;
    FCHK(q, hipMemcpyAsync(q->h_rings.data(), p.rings, sizeof(double) * 3 * R,
                           hipMemcpyDeviceToHost, q->stream));
    FCHK(q, hipStreamSynchronize(q->stream));
    const double res = frc_resolution(q->h_rings.data(), p.N, curve);
    if (resolution) *resolution = res;
    if (n_rings) *n_rings = R;
    return ESRRF_OK;
}

}  // namespace

int esrrf_frc_create(int width, int height, int device, esrrf_frc **out)
{
    if (!out || width < 2 || height < 2) return ESRRF_ERR_INVALID;
    *out = nullptr;
    const int S = std::min(width, height);
    int N = 2, lg = 1;
    while (N < S) {
        N <<= 1;
        lg++;
    }
    if (N > FRC_MAX_N) return ESRRF_ERR_UNSUPPORTED;
    esrrf_frc *q = new esrrf_frc();
    auto bad = [&](int code) {
        (void)hipGetLastError();   // a failed call must not surface in later checks
        frc_free(q);
        delete q;
        return code;
    };
    int dev = device;
    if (dev < 0 && esrrf_fastest_device(&dev) != ESRRF_OK) return bad(ESRRF_ERR_HIP);
    q->dev = dev;
    if (hipSetDevice(dev) != hipSuccess) return bad(ESRRF_ERR_HIP);
    if (hipStreamCreate(&q->own_stream) != hipSuccess) return bad(ESRRF_ERR_HIP);
    q->stream = q->own_stream;
    FrcPlan &p = q->plan;
    p.W = width;
    p.H = height;
    p.S = S;
    p.x0 = (width - S) / 2;
    p.y0 = (height - S) / 2;
    p.N = N;
    p.log2N = lg;
    p.nparts = 512;
    const int R = N / 2;
    std::vector<float> w = tukey(S);
    std::vector<float2> tw(N);
    for (int k = 0; k < N; k++) {
        const double ang = -2.0 * M_PI * k / N;
        tw[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    float *d_win = nullptr;
    float2 *d_tw = nullptr;
    if (hipMalloc((void **)&d_win, sizeof(float) * S) != hipSuccess) return bad(ESRRF_ERR_NOMEM);
    p.win = d_win;
    if (hipMalloc((void **)&d_tw, sizeof(float2) * N) != hipSuccess) return bad(ESRRF_ERR_NOMEM);
    p.tw = d_tw;
    if (hipMalloc((void **)&p.buf, sizeof(float2) * 2 * (size_t)N * N) != hipSuccess ||
        hipMalloc((void **)&p.part, sizeof(double) * 3 * (size_t)R * p.nparts) != hipSuccess ||
        hipMalloc((void **)&p.rings, sizeof(double) * 3 * (size_t)R) != hipSuccess)
        return bad(ESRRF_ERR_NOMEM);
    if (hipMemcpy(d_win, w.data(), sizeof(float) * S, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_tw, tw.data(), sizeof(float2) * N, hipMemcpyHostToDevice) != hipSuccess)
        return bad(ESRRF_ERR_HIP);
    q->h_rings.resize(3 * (size_t)R);
    *out = q;
    return ESRRF_OK;
}

int esrrf_frc_destroy(esrrf_frc *q)
{
    if (!q) return ESRRF_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(q->mu);
        (void)hipSetDevice(q->dev);
        (void)hipStreamSynchronize(q->stream);
        frc_free(q);
    }
    delete q;
    return ESRRF_OK;
}

int esrrf_frc_set_stream(esrrf_frc *q, void *hip_stream)
{
    if (!q) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(q->mu);
    q->stream = hip_stream ? (hipStream_t)hip_stream : q->own_stream;
    return ESRRF_OK;
}

int esrrf_frc_rings(esrrf_frc *q, int *n_rings, int *transform_size)
{
    if (!q) return ESRRF_ERR_INVALID;
    if (n_rings) *n_rings = q->plan.N / 2;
    if (transform_size) *transform_size = q->plan.N;
    return ESRRF_OK;
}

int esrrf_frc_run_device(esrrf_frc *q, const float *d_a, const float *d_b, double *resolution,
                         float *curve)
{
    if (!q) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(q->mu);
    if (!d_a || !d_b) return ffail(q, ESRRF_ERR_INVALID, "both images must be given");
    FCHK(q, hipSetDevice(q->dev));
    return run_impl(q, d_a, d_b, resolution, curve, nullptr);
}

int esrrf_frc_run(esrrf_frc *q, const float *a, const float *b, double *resolution,
                  float *curve)
{
    if (!q) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(q->mu);
    if (!a || !b) return ffail(q, ESRRF_ERR_INVALID, "both images must be given");
    FCHK(q, hipSetDevice(q->dev));
    const size_t n = (size_t)q->plan.W * q->plan.H;
    if (!q->d_stage) FCHK(q, hipMalloc((void **)&q->d_stage, sizeof(float) * 2 * n));
    FCHK(q, hipMemcpyAsync(q->d_stage, a, sizeof(float) * n, hipMemcpyHostToDevice, q->stream));
    FCHK(q, hipMemcpyAsync(q->d_stage + n, b, sizeof(float) * n, hipMemcpyHostToDevice,
                           q->stream));
    return run_impl(q, q->d_stage, q->d_stage + n, resolution, curve, nullptr);
}

const char *esrrf_frc_last_error(esrrf_frc *q) { return q ? q->err.c_str() : ""; }
