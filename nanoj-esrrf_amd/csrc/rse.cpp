// rse.cpp -- SQUIRREL RSE of the eSRRF AVG image (ErrorMapLiveSRRF.java:81-246, as the
// parameter sweep uses it: ref = INT, SR = AVG after MP correction, gui/ParametersSweep_.java:473)
// and its gradient with respect to the continuous parameters (fwhm, sensitivity, RSF sigma).
//
// Gradient: alpha and beta are the least-squares fit of ref on (blur(SR), blur(1)), so by
// the envelope theorem dE/dtheta = -2 alpha sum r d(blur(SR))/dtheta at fixed alpha, beta
// (r = ref - alpha blur(SR) - beta, E = sum r^2, RSE = sqrt(E/n)).  d(SR)/d(fwhm, S) come
// from the adjoint variant of the tiled RGC kernel, which recomputes every frame's CGLH and
// its tSS-derivative per block and accumulates dAVG/dtSS and dAVG/dS beside AVG (no
// per-frame storage); the MP correction is differentiated in forward mode; d/dsigma is the
// blur with the sigma-derivative of its taps.  The window size and the d <= tSO gate are
// piecewise constant in fwhm and are held fixed by the derivative.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/esrrf.h"
#include "bic_launch.h"
#include "esrrf_internal.h"
#include "rse_internal.h"

using namespace esrrf;

struct esrrf_rse {
    std::mutex mu;
    std::string err;
    int dev = 0;
    hipStream_t own_stream = nullptr, stream = nullptr;
    int w = 0, h = 0, M = 1, nF = 0, crop_raw = 0, mp = 1, intw = 1;
    float sigma_max = 0.0f;
    void *d_in = nullptr;
    float *d_px = nullptr;
    float2 *d_U = nullptr;
    float *d_out4 = nullptr, *d_prev = nullptr;
    int loaded = 0;
    BicTables tabs;
    float tabs_fwhm = -1.0f;
    BicShape shape;          // adjoint launches (mode 2)
    BicShape shape_value;    // value-only launches (mode 0)
    RsePlan plan{};
    // (fwhm, S) of the current SR image in plan.SR (and of Dt/DS when cur_grad)
    float cur_fwhm = -1.0f, cur_sens = -1.0f;
    bool cur_grad = false;
    long long evals = 0, rgc_passes = 0;
};

namespace {

int rfail(esrrf_rse *r, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (r) r->err = buf;
    return code;
}

#define RCHK(r, x)                                                                         \
    do {                                                                                   \
        hipError_t _st = (x);                                                              \
        if (_st != hipSuccess)                                                             \
            return rfail((r), ESRRF_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(_st)); \
    } while (0)

void rse_free(esrrf_rse *r)
{
    RsePlan &p = r->plan;
    void *ptrs[] = {r->d_in, r->d_px, r->d_U, r->d_out4, r->d_prev, p.stats, p.Y, p.SR, p.Dt,
                    p.DS, p.T1, p.T2, p.X, p.G1, p.G2, p.G3, p.taps, p.part, p.red};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    bic_tables_free(r->tabs);
    if (r->own_stream) (void)hipStreamDestroy(r->own_stream);
}

// The cropped MP-corrected SR image of (fwhm, S) over the loaded frames: with grad, from the
// adjoint launch (AVG, dAVG/dtSS, dAVG/dS, INT) and with its derivatives; without, from the
// plain cumulant launch (AVG, E2, TAC2, INT), as the reference computes a configuration.
int prepare_sr(esrrf_rse *r, float fwhm, float sens, bool grad)
{
    if (r->cur_fwhm == fwhm && r->cur_sens == sens && (r->cur_grad || !grad)) return ESRRF_OK;
    if (r->tabs_fwhm != fwhm) {
        bic_tables_free(r->tabs);
        r->tabs_fwhm = -1.0f;
        RCHK(r, bic_tables_create(r->w, r->h, r->M, fwhm, r->tabs));
        if (!r->tabs.ok) return rfail(r, ESRRF_ERR_UNSUPPORTED, "fwhm %g: window too large", fwhm);
        r->tabs_fwhm = fwhm;
        r->shape = bic_default_shape(r->tabs, r->M, true);        // adjoint variant
        r->shape_value = bic_default_shape(r->tabs, r->M, false);  // value only
    }
    const BicShape &sh = grad ? r->shape : r->shape_value;
    BicRgcArgs a = bic_args(r->tabs, sh, r->d_U, r->d_px, r->w, r->h, r->M);
    a.mode = grad ? 2 : 0;
    a.out = r->d_out4;
    a.prev = r->d_prev;
    a.frame0 = 0;
    a.nframes = r->loaded;
    a.first_init = 1;
    a.rn = 1.0f / (float)r->loaded;
    a.rn1 = 1.0f / (float)(r->loaded - 1);
    a.sens = sens;
    a.sens_int = sens_exact_int(sens);
    a.int_weighting = r->intw;
    RCHK(r, launch_bic_rgc(a, sh.ppw, r->stream));
    RCHK(r, launch_rse_prep(r->plan, r->d_out4, r->stream));
    r->cur_fwhm = fwhm;
    r->cur_sens = sens;
    r->cur_grad = grad;
    r->rgc_passes++;
    return ESRRF_OK;
}

int eval_impl(esrrf_rse *r, float fwhm, float sens, float sigma, double *rse, double *grad)
{
    if (r->loaded < 2) return rfail(r, ESRRF_ERR_STATE, "load at least 2 frames first");
    if (!(fwhm > 0.0f) || !(sens > 0.0f) || !(sigma >= 0.0f) || sigma > r->sigma_max)
        return rfail(r, ESRRF_ERR_INVALID, "parameters outside fwhm > 0, S > 0, 0 <= sigma <= %g",
                     r->sigma_max);
    int rc = prepare_sr(r, fwhm, sens, grad != nullptr);
    if (rc) return rc;
    RCHK(r, launch_rse_eval(r->plan, sigma, grad != nullptr, r->stream));
    double red[10];
    RCHK(r, hipMemcpyAsync(red, r->plan.red, sizeof red, hipMemcpyDeviceToHost, r->stream));
    RCHK(r, hipStreamSynchronize(r->stream));
    r->evals++;
    const double n = (double)r->plan.n;
    const double E = std::max(red[6], 0.0);
    const double R = std::sqrt(E / n);
    if (rse) *rse = R;
    if (grad) {
        const double alpha = red[4];
        const double k = R > 0.0 ? -alpha / (n * R) : 0.0;
        // tSS = 2 (fwhm / 2.354)^2 (LiveSRRF_CL.java:304,318)
        const double dtss = 4.0 * (double)fwhm / (2.354 * 2.354);
        grad[0] = k * red[7] * dtss;
        grad[1] = k * red[8];
        grad[2] = k * red[9];
    }
    return ESRRF_OK;
}

}  // namespace

int esrrf_rse_create(const esrrf_rse_params *p, esrrf_rse **out)
{
    if (!p || !out) return ESRRF_ERR_INVALID;
    *out = nullptr;
    if (p->width < 4 || p->height < 4 || p->magnification < 1 || p->magnification > 8 ||
        p->n_frames < 2 || p->crop_size < 0 || 2 * p->crop_size >= std::min(p->width, p->height))
        return ESRRF_ERR_INVALID;
    esrrf_rse *r = new esrrf_rse();
    auto bad = [&](int code) {
        (void)hipGetLastError();   // a failed call must not surface in later checks
        rse_free(r);
        delete r;
        return code;
    };
    r->w = p->width;
    r->h = p->height;
    r->M = p->magnification;
    r->nF = p->n_frames;
    r->crop_raw = p->crop_size;
    r->mp = p->do_mp_map_correction ? 1 : 0;
    r->intw = p->int_weighting ? 1 : 0;
    // ErrorMapLiveSRRF.java:75: 4/2.35482 raw pixels (Nyquist) in SR pixels
    r->sigma_max = p->sigma_max > 0.0f ? p->sigma_max : (4.0f / 2.35482f) * (float)r->M;
    int dev = p->device;
    if (dev < 0 && esrrf_fastest_device(&dev) != ESRRF_OK) return bad(ESRRF_ERR_HIP);
    r->dev = dev;
    if (hipSetDevice(dev) != hipSuccess) return bad(ESRRF_ERR_HIP);
    if (hipStreamCreate(&r->own_stream) != hipSuccess) return bad(ESRRF_ERR_HIP);
    r->stream = r->own_stream;
    const long wh = (long)r->w * r->h;
    const long whM = wh * r->M * r->M;
    RsePlan &q = r->plan;
    q.w = r->w;
    q.h = r->h;
    q.M = r->M;
    q.crop = r->crop_raw * r->M;
    q.Wc = r->w * r->M - 2 * q.crop;
    q.Hc = r->h * r->M - 2 * q.crop;
    q.n = (long)q.Wc * q.Hc;
    q.mp = r->mp;
    q.K = std::max(1, (int)std::ceil(5.0 * (double)r->sigma_max));
    auto alloc = [&](void **ptr, size_t bytes) { return hipMalloc(ptr, bytes ? bytes : 8) == hipSuccess; };
    const size_t cb = sizeof(float) * q.n;
    bool ok = alloc(&r->d_in, sizeof(float) * wh * r->nF) &&
              alloc((void **)&r->d_px, sizeof(float) * wh * r->nF) &&
              alloc((void **)&r->d_U, sizeof(float2) * 4 * wh * r->nF) &&
              alloc((void **)&r->d_out4, sizeof(float) * 4 * whM) &&
              alloc((void **)&r->d_prev, sizeof(float) * whM) &&
              alloc((void **)&q.stats, sizeof(double) * 3 * r->M * r->M);
    float **bufs[] = {&q.Y, &q.SR, &q.Dt, &q.DS, &q.T1, &q.T2, &q.X, &q.G1, &q.G2, &q.G3};
    for (float **b : bufs) ok = ok && alloc((void **)b, cb);
    ok = ok && alloc((void **)&q.taps, sizeof(float) * 2 * (2 * q.K + 1)) &&
         alloc((void **)&q.part, sizeof(double) * 4 * rse_max_parts(q.n)) &&
         alloc((void **)&q.red, sizeof(double) * 12);
    if (!ok) return bad(ESRRF_ERR_NOMEM);
    if (hipMemset(q.stats, 0, sizeof(double) * 3 * r->M * r->M) != hipSuccess)
        return bad(ESRRF_ERR_HIP);
    *out = r;
    return ESRRF_OK;
}

int esrrf_rse_destroy(esrrf_rse *r)
{
    if (!r) return ESRRF_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(r->mu);
        (void)hipSetDevice(r->dev);
        (void)hipStreamSynchronize(r->stream);
        rse_free(r);
    }
    delete r;
    return ESRRF_OK;
}

int esrrf_rse_load(esrrf_rse *r, const void *frames, int dtype, int n_frames)
{
    if (!r || !frames) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(r->mu);
    if (dtype != ESRRF_U16 && dtype != ESRRF_F32) return rfail(r, ESRRF_ERR_INVALID, "dtype");
    if (n_frames < 2 || n_frames > r->nF)
        return rfail(r, ESRRF_ERR_INVALID, "2 <= n_frames <= %d", r->nF);
    (void)hipSetDevice(r->dev);
    const long wh = (long)r->w * r->h;
    RCHK(r, hipMemcpyAsync(r->d_in, frames, (size_t)wh * n_frames * (dtype == ESRRF_U16 ? 2 : 4),
                           hipMemcpyHostToDevice, r->stream));
    RCHK(r, launch_convert_frames(r->d_in, dtype, r->d_px, wh * n_frames, r->stream));
    RCHK(r, launch_bic_prep(r->d_px, r->d_U, r->w, r->h, n_frames, r->stream));
    RCHK(r, hipStreamSynchronize(r->stream));
    r->loaded = n_frames;
    r->cur_fwhm = r->cur_sens = -1.0f;
    return ESRRF_OK;
}

int esrrf_rse_value_and_grad(esrrf_rse *r, float fwhm, float sensitivity, float sigma,
                             double *rse, double *grad3)
{
    if (!r) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(r->mu);
    (void)hipSetDevice(r->dev);
    return eval_impl(r, fwhm, sensitivity, sigma, rse, grad3);
}

int esrrf_rse_best_sigma(esrrf_rse *r, float fwhm, float sensitivity, double *sigma, double *rse,
                         int *n_eval)
{
    if (!r) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(r->mu);
    (void)hipSetDevice(r->dev);
    auto f = [&](double s, double *v) {
        return eval_impl(r, fwhm, sensitivity, (float)s, v, nullptr);
    };
    double xs = 0.0, fs = 0.0;
    int rc = brent(f, 0.0, (double)r->sigma_max, &xs, &fs, n_eval);
    if (rc) return rc;
    if (sigma) *sigma = xs;
    if (rse) *rse = fs;
    return ESRRF_OK;
}

int esrrf_rse_descend(esrrf_rse *r, const float *x0, int steps, const float *lr, const float *lo,
                      const float *hi, double *trajectory)
{
    if (!r || !x0 || steps < 0) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(r->mu);
    (void)hipSetDevice(r->dev);
    // Adam (Kingma and Ba, 2015) on (fwhm, sensitivity, sigma), projected on the box
    const double dlr[3] = {0.1, 0.25, 0.5};
    const double dlo[3] = {0.3, 0.25, 0.01};
    const double dhi[3] = {8.0, 10.0, (double)r->sigma_max};
    double x[3], m[3] = {0, 0, 0}, v[3] = {0, 0, 0}, step[3], l[3], u[3];
    for (int i = 0; i < 3; i++) {
        step[i] = lr ? lr[i] : dlr[i];
        l[i] = lo ? lo[i] : dlo[i];
        u[i] = hi ? std::min((double)hi[i], i == 2 ? (double)r->sigma_max : 1e30) : dhi[i];
        x[i] = std::min(std::max((double)x0[i], l[i]), u[i]);
    }
    const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    for (int k = 0; k <= steps; k++) {
        double R, g[3];
        int rc = eval_impl(r, (float)x[0], (float)x[1], (float)x[2], &R, k < steps ? g : nullptr);
        if (rc) return rc;
        if (trajectory) {
            trajectory[4 * k + 0] = (float)x[0];
            trajectory[4 * k + 1] = (float)x[1];
            trajectory[4 * k + 2] = (float)x[2];
            trajectory[4 * k + 3] = R;
        }
        if (k == steps) break;
        for (int i = 0; i < 3; i++) {
            if (!std::isfinite(g[i])) g[i] = 0.0;
            m[i] = b1 * m[i] + (1 - b1) * g[i];
            v[i] = b2 * v[i] + (1 - b2) * g[i] * g[i];
            const double mh = m[i] / (1 - std::pow(b1, k + 1));
            const double vh = v[i] / (1 - std::pow(b2, k + 1));
            x[i] = std::min(std::max(x[i] - step[i] * mh / (std::sqrt(vh) + eps), l[i]), u[i]);
        }
    }
    return ESRRF_OK;
}

int esrrf_rse_counters(esrrf_rse *r, long long *evaluations, long long *rgc_passes)
{
    if (!r) return ESRRF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(r->mu);
    if (evaluations) *evaluations = r->evals;
    if (rgc_passes) *rgc_passes = r->rgc_passes;
    return ESRRF_OK;
}

const char *esrrf_rse_last_error(esrrf_rse *r) { return r ? r->err.c_str() : "null rse"; }
