// squirrel.cpp -- SQUIRREL error map of any (reference, SR) image pair
// (ErrorMapLiveSRRF.java:50-193) and the QnR map (gui/GetQnRmap_.java:179-260).
//
// optimise(impRef, impSR) + calculateErrorMap() as one call: crop, Brent search of the RSF
// sigma on [0, 4/2.35482 M] (or the fixed sigma), alpha/beta at that sigma, then RSC, the
// error map, RSE and RSP.  Each objective evaluation is the blur / alpha-beta / residual
// sequence of the RSE pipeline and one read-back of its sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/esrrf.h"
#include "rse_internal.h"

using namespace esrrf;

struct esrrf_squirrel {
    std::mutex mu;
    std::string err;
    int dev = 0;
    hipStream_t own_stream = nullptr, stream = nullptr;
    int w = 0, h = 0, M = 1, crop = 0, fix_sigma = 0;
    float fixed_sigma = 0.0f, sigma_max = 0.0f;
    RsePlan plan{};
    double *d_part = nullptr;   // k_sq_maps partial sums
    double *d_out = nullptr;    // RSE, RSP, alpha, beta
    float *d_full = nullptr;    // host-call staging: ref and SR [2][hM][wM]
    float *d_maps = nullptr;    // host-call staging: error map and RSC [2][Hc][Wc]
};

namespace {

int qfail(esrrf_squirrel *q, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (q) q->err = buf;
    return code;
}

#define QCHK(q, x)                                                                         \
    do {                                                                                   \
        hipError_t _st = (x);                                                              \
        if (_st != hipSuccess)                                                             \
            return qfail((q), ESRRF_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(_st)); \
    } while (0)

void squirrel_free(esrrf_squirrel *q)
{
    RsePlan &p = q->plan;
    void *ptrs[] = {p.Y, p.SR, p.T1, p.X, p.taps, p.part, p.red, q->d_part, q->d_out,
                    q->d_full, q->d_maps};
    for (void *x : ptrs)
        if (x) (void)hipFree(x);
    if (q->own_stream) (void)hipStreamDestroy(q->own_stream);
}

// RSE(sigma) of the cropped pair in plan.Y / plan.SR (sigmaOptimiseFunction.value, :262-287)
int objective(esrrf_squirrel *q, double sigma, double *rse)
{
    QCHK(q, launch_rse_eval(q->plan, (float)sigma, false, q->stream));
    double red[10];
    QCHK(q, hipMemcpyAsync(red, q->plan.red, sizeof red, hipMemcpyDeviceToHost, q->stream));
    QCHK(q, hipStreamSynchronize(q->stream));
    *rse = std::sqrt(std::max(red[6], 0.0) / (double)q->plan.n);
    return ESRRF_OK;
}

int run_impl(esrrf_squirrel *q, const float *d_ref, const float *d_sr, esrrf_squirrel_result *res,
             float *d_emap, float *d_rsc)
{
    RsePlan &p = q->plan;
    QCHK(q, launch_squirrel_crop(d_ref, d_sr, q->w * q->M, p.crop, p, q->stream));
    double sigma = 0.0, best = 0.0;
    int nev = 0;
    if (q->fix_sigma) {
        sigma = (double)q->fixed_sigma;
    } else {
        auto f = [&](double s, double *v) { return objective(q, s, v); };
        const int rc = brent(f, 0.0, (double)q->sigma_max, &sigma, &best, &nev);
        if (rc) return rc;
    }
    // sigma_linear is a float (:127); alpha, beta and the maps at that sigma
    const float sig = (float)sigma;
    QCHK(q, launch_rse_eval(p, sig, false, q->stream));
    QCHK(q, launch_squirrel_maps(p, d_emap, d_rsc, q->d_part, q->d_out, q->stream));
    double out[4];
    QCHK(q, hipMemcpyAsync(out, q->d_out, sizeof out, hipMemcpyDeviceToHost, q->stream));
    QCHK(q, hipStreamSynchronize(q->stream));
    if (res) {
        res->sigma = (double)sig;
        res->rse = out[0];
        res->rsp = out[1];
        res->alpha = out[2];
        res->beta = out[3];
        res->overblur = std::fabs(sig - q->sigma_max) < 0.0001f ? 1 : 0;
        res->evaluations = nev;
    }
    return ESRRF_OK;
}

}  // namespace

int esrrf_squirrel_create(const esrrf_squirrel_params *pp, esrrf_squirrel **out)
{
    if (!pp || !out) return ESRRF_ERR_INVALID;
    *out = nullptr;
    const int crop_raw = pp->crop_border ? pp->crop_size : 0;
    if (pp->width < 1 || pp->height < 1 || pp->magnification < 1 || pp->magnification > 16 ||
        crop_raw < 0 || 2 * crop_raw >= std::min(pp->width, pp->height) ||
        (pp->fix_sigma && !(pp->fixed_sigma >= 0.0f)))
        return ESRRF_ERR_INVALID;
    esrrf_squirrel *q = new esrrf_squirrel();
    auto bad = [&](int code) {
        (void)hipGetLastError();   // a failed call must not surface in later checks
        squirrel_free(q);
        delete q;
        return code;
    };
    q->w = pp->width;
    q->h = pp->height;
    q->M = pp->magnification;
    q->fix_sigma = pp->fix_sigma ? 1 : 0;
    q->fixed_sigma = pp->fixed_sigma * (float)q->M;   // fixedSigmaSRspace (:53)
    q->sigma_max = (4.0f / 2.35482f) * (float)q->M;    // maxSigmaBoundary (:75)
    int dev = pp->device;
    if (dev < 0 && esrrf_fastest_device(&dev) != ESRRF_OK) return bad(ESRRF_ERR_HIP);
    q->dev = dev;
    if (hipSetDevice(dev) != hipSuccess) return bad(ESRRF_ERR_HIP);
    if (hipStreamCreate(&q->own_stream) != hipSuccess) return bad(ESRRF_ERR_HIP);
    q->stream = q->own_stream;
    RsePlan &p = q->plan;
    p.w = q->w;
    p.h = q->h;
    p.M = q->M;
    p.crop = crop_raw * q->M;
    p.Wc = q->w * q->M - 2 * p.crop;
    p.Hc = q->h * q->M - 2 * p.crop;
    p.n = (long)p.Wc * p.Hc;
    p.mp = 0;
    const double smax = std::max((double)q->sigma_max, q->fix_sigma ? (double)q->fixed_sigma : 0.0);
    p.K = std::max(1, (int)std::ceil(5.0 * smax));
    auto alloc = [&](void **ptr, size_t bytes) {
        return hipMalloc(ptr, bytes ? bytes : 8) == hipSuccess;
    };
    const size_t cb = sizeof(float) * p.n;
    bool ok = true;
    float **bufs[] = {&p.Y, &p.SR, &p.T1, &p.X};
    for (float **b : bufs) ok = ok && alloc((void **)b, cb);
    ok = ok && alloc((void **)&p.taps, sizeof(float) * 2 * (2 * p.K + 1)) &&
         alloc((void **)&p.part, sizeof(double) * 4 * rse_max_parts(p.n)) &&
         alloc((void **)&p.red, sizeof(double) * 12) &&
         alloc((void **)&q->d_part, sizeof(double) * 6 * squirrel_max_parts(p.n)) &&
         alloc((void **)&q->d_out, sizeof(double) * 4);
    if (!ok) return bad(ESRRF_ERR_NOMEM);
    *out = q;
    return ESRRF_OK;
}

int esrrf_squirrel_destroy(esrrf_squirrel *q)
{
    if (!q) return ESRRF_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(q->mu);
        (void)hipSetDevice(q->dev);
        (void)hipStreamSynchronize(q->stream);
        squirrel_free(q);
    }
    delete q;
    return ESRRF_OK;
}

int esrrf_squirrel_set_stream(esrrf_squirrel *q, void *hip_stream)
{
    if (!q) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(q->mu);
    q->stream = hip_stream ? (hipStream_t)hip_stream : q->own_stream;
    return ESRRF_OK;
}

int esrrf_squirrel_dims(esrrf_squirrel *q, int *width_c, int *height_c)
{
    if (!q) return ESRRF_ERR_INVALID;
    if (width_c) *width_c = q->plan.Wc;
    if (height_c) *height_c = q->plan.Hc;
    return ESRRF_OK;
}

int esrrf_squirrel_run_device(esrrf_squirrel *q, const float *d_ref, const float *d_sr,
                              esrrf_squirrel_result *res, float *d_error_map, float *d_rsc)
{
    if (!q) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(q->mu);
    if (!d_ref || !d_sr) return qfail(q, ESRRF_ERR_INVALID, "ref and sr must be given");
    QCHK(q, hipSetDevice(q->dev));
    return run_impl(q, d_ref, d_sr, res, d_error_map, d_rsc);
}

int esrrf_squirrel_run(esrrf_squirrel *q, const float *ref, const float *sr,
                       esrrf_squirrel_result *res, float *error_map, float *rsc)
{
    if (!q) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(q->mu);
    if (!ref || !sr) return qfail(q, ESRRF_ERR_INVALID, "ref and sr must be given");
    QCHK(q, hipSetDevice(q->dev));
    const long full = (long)q->w * q->M * q->h * q->M;
    const long n = q->plan.n;
    if (!q->d_full) QCHK(q, hipMalloc((void **)&q->d_full, sizeof(float) * 2 * full));
    if (!q->d_maps) QCHK(q, hipMalloc((void **)&q->d_maps, sizeof(float) * 2 * n));
    QCHK(q, hipMemcpyAsync(q->d_full, ref, sizeof(float) * full, hipMemcpyHostToDevice, q->stream));
    QCHK(q, hipMemcpyAsync(q->d_full + full, sr, sizeof(float) * full, hipMemcpyHostToDevice,
                           q->stream));
    const int rc = run_impl(q, q->d_full, q->d_full + full, res, error_map ? q->d_maps : nullptr,
                            rsc ? q->d_maps + n : nullptr);
    if (rc) return rc;
    if (error_map)
        QCHK(q, hipMemcpyAsync(error_map, q->d_maps, sizeof(float) * n, hipMemcpyDeviceToHost,
                               q->stream));
    if (rsc)
        QCHK(q, hipMemcpyAsync(rsc, q->d_maps + n, sizeof(float) * n, hipMemcpyDeviceToHost,
                               q->stream));
    QCHK(q, hipStreamSynchronize(q->stream));
    return ESRRF_OK;
}

const char *esrrf_squirrel_last_error(esrrf_squirrel *q) { return q ? q->err.c_str() : ""; }

int esrrf_remove_nans(float *map, int width, int height, double radius)
{
    if (!map || width < 1 || height < 1 || !(radius >= 0.0) || radius > 64.0)
        return ESRRF_ERR_INVALID;
    // kernel lines: for dy in [-kr, kr], dx in [-dxmax(dy), dxmax(dy)] with
    // dx^2 + dy^2 <= r^2 + 1
    const double r2 = radius * radius + 1.0;
    const int kr = (int)std::sqrt(r2 + 1e-10);
    std::vector<int> dxmax(2 * kr + 1);
    for (int dy = -kr; dy <= kr; dy++) dxmax[dy + kr] = (int)std::sqrt(r2 - dy * dy + 1e-10);
    const std::vector<float> src(map, map + (size_t)width * height);
    std::vector<float> vals;
    for (int y = 0; y < height; y++)
        for (int x = 0; x < width; x++) {
            if (!std::isnan(src[(size_t)y * width + x])) continue;
            vals.clear();
            for (int dy = -kr; dy <= kr; dy++) {
                const int yy = std::min(std::max(y + dy, 0), height - 1);
                for (int dx = -dxmax[dy + kr]; dx <= dxmax[dy + kr]; dx++) {
                    const int xx = std::min(std::max(x + dx, 0), width - 1);
                    const float v = src[(size_t)yy * width + xx];
                    if (!std::isnan(v)) vals.push_back(v);
                }
            }
            if (vals.empty()) continue;
            std::sort(vals.begin(), vals.end());
            const size_t n = vals.size();
            map[(size_t)y * width + x] =
                n & 1 ? vals[n / 2] : (float)(0.5 * ((double)vals[n / 2 - 1] + vals[n / 2]));
        }
    return ESRRF_OK;
}

int esrrf_qnr_map(const float *frc, const float *rsp, long n, int auto_min_max, float res_min,
                  float res_max, float *qnr, float *best, long *best_index)
{
    if (!frc || !rsp || !qnr || n < 1) return ESRRF_ERR_INVALID;
    if (auto_min_max) {
        // getMinMaxfromImageStack (gui/GetQnRmap_.java:290-305): finite values only
        float mn = 3.4028235e38f, mx = -3.4028235e38f;
        for (long i = 0; i < n; i++) {
            const float v = frc[i];
            if (!std::isfinite(v)) continue;
            mn = std::min(mn, v);
            mx = std::max(mx, v);
        }
        res_min = mn;
        res_max = mx;
    }
    if (!(res_max != res_min)) return ESRRF_ERR_INVALID;
    // logisticImageConversion (:179-203), in double as written
    const double M1 = 0.075, M2 = 0.925;
    const double A1 = std::log((1.0 - M1) / M1), A2 = std::log((1.0 - M2) / M2);
    const double x0 = (A2 * (double)res_max - A1 * (double)res_min) / (A2 - A1);
    const double k = 1.0 / (x0 - (double)res_max) * A1;
    float bv = -3.4028235e38f;
    long bi = 0;
    for (long i = 0; i < n; i++) {
        const float L = (float)(1.0 / (std::exp(-k * ((double)frc[i] - x0)) + 1.0));
        // calculateQnRmap (:227-234)
        const float q = 2.0f * L * rsp[i] / (L + rsp[i]);
        qnr[i] = q;
        if (q > bv) {   // findMaximum (:244-258): first strict maximum
            bv = q;
            bi = i;
        }
    }
    if (best) *best = bv;
    if (best_index) *best_index = bi;
    return ESRRF_OK;
}
