// modes.cpp -- host side of the FHT (Fourier-upsampled) and 3-D (multi-focus) pipelines.
//
//  FHT : LiveSRRF_CL.prepareBuffersFHTinterpolation (LiveSRRF_CL.java:462-579),
//        prepareDataFHT_SRRF (:598-632) and the RGC loop of calculateSRRF (:699-746);
//        MP correction is never applied in this mode (:762).
//  3-D : prepareBuffersBICinterpolation with live3DSRRF.cl (:266-459), the montage split
//        of reshapeImageStack3D (:963-994) and calculateSRRF's 3-D branch (:655-696).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "esrrf_internal.h"
#include "fht_internal.h"
#include "modes.h"
#include "mode3d_internal.h"

#pragma STDC FP_CONTRACT OFF

namespace esrrf {

namespace {

int err(ModeState &ms, int code, const char *msg, hipError_t st = hipSuccess)
{
    ms.err = msg;
    if (st != hipSuccess) {
        ms.err += ": ";
        ms.err += hipGetErrorString(st);
    }
    return code;
}

template <class T>
bool dmalloc(T **p, size_t bytes)
{
    *p = nullptr;
    if (hipMalloc((void **)p, bytes ? bytes : 4) == hipSuccess) return true;
    (void)hipGetLastError();   // a failed allocation must not surface in later checks
    return false;
}

// Java: (int) Math.ceil(Math.log(n)/Math.log(2)) then (int) Math.pow(2, .)
int pow2_ceil_java(int n)
{
    const int e = (int)std::ceil(std::log((double)n) / std::log(2.0));
    return (int)std::pow(2.0, (double)e);
}

// Tap table of the FHT-mode RGC for unshifted frames: the window offsets (i, j) in the
// reference's j-then-i order (liveSRRF_noInterpolation.cl:276-289), dx = i, dy = j exactly.
void build_fht_taps(ModeState &ms, float radius, float tSS, float tSO)
{
    const int R = (int)radius, Rhi = (int)(radius + 1.0f);
    const int TW = 16 + R + Rhi;
    int pitch = 8;
    while (pitch < TW) pitch += 32;
    ms.fht_R = R;
    ms.fht_Rhi = Rhi;
    ms.fht_TW = TW;
    ms.fht_pitch = pitch;
    ms.h_taps.clear();
    ms.h_tapk.clear();
    float wsum = 0.0f;
    for (int j = -R; j <= Rhi; j++) {
        for (int i = -R; i <= Rhi; i++) {
            volatile float dx = (float)i, dy = (float)j;
            volatile float d2 = dx * dx;
            volatile float dy2 = dy * dy;
            d2 = d2 + dy2;
            const float d = std::sqrt((float)d2);
            if (!(d != 0.0f && d <= tSO)) continue;
            volatile float e = -(d * d);
            e = e / tSS;
            volatile float w1 = d * std::exp((float)e);
            volatile float W = w1 * w1;
            W = W * w1;
            W = W * w1;
            wsum = wsum + (float)W;
            Tap t;
            t.off = j * pitch + i;
            t.a = (float)dx / d;
            t.b = (float)dy / d;
            t.W = (float)W;
            ms.h_taps.push_back(t);
            ms.h_tapk.push_back((i + 128) | ((j + 128) << 8));
        }
    }
    ms.ntaps = (int)ms.h_taps.size();
    ms.tap_wsum = wsum;
    while (ms.h_taps.size() % 4 != 0 || ms.h_taps.size() < (size_t)ms.ntaps + 4) {
        Tap z;
        z.off = 0;
        z.a = z.b = z.W = 0.0f;
        ms.h_taps.push_back(z);
        ms.h_tapk.push_back(128 | (128 << 8));
    }
}

}  // namespace

int mode_create(const ModeCreate &c, ModeState &ms, float **d_px, float2 **d_U, float **d_aux,
                size_t *aux_bytes)
{
    ms.mode = c.mode;
    ms.nF = c.nF;
    ms.var = c.var;
    *aux_bytes = 0;
    *d_aux = nullptr;
    if (c.mode == 1) {
        const long wM = (long)c.w * c.M, hM = (long)c.h * c.M;
        if (!dmalloc(d_px, sizeof(float) * wM * hM * c.nF) ||
            !dmalloc(d_U, sizeof(float2) * wM * hM * c.nF) ||
            !dmalloc(&ms.d_rawf, sizeof(float) * (size_t)c.w * c.h * c.nF))
            return err(ms, ESRRF_ERR_NOMEM, "FHT buffers: hipMalloc failed");
        if (c.M > 1) {
            // mirror padding to a square power of two (FourierInterpolation.java:118-172)
            const int bw = pow2_ceil_java(c.w), bh = pow2_ceil_java(c.h);
            const int N = std::max(bw, bh);
            if (N > 4096) return err(ms, ESRRF_ERR_UNSUPPORTED, "FHT: frames larger than 4096");
            int sW = (int)std::ceil((double)N / c.w);
            if (sW % 2 == 0) sW += 1;
            int sH = (int)std::ceil((double)N / c.h);
            if (sH % 2 == 0) sH += 1;
            FhtPlan &p = ms.plan;
            p.N = N;
            p.log2N = 0;
            while ((1 << p.log2N) < N) p.log2N++;
            p.M = c.M;
            p.w = c.w;
            p.h = c.h;
            p.scaleW = sW;
            p.scaleH = sH;
            p.xROIp = (sW * c.w - N) / 2;
            p.yROIp = (sH * c.h - N) / 2;
            p.xROI = (c.M * N - c.M * c.w) / 2;
            p.yROI = (c.M * N - c.M * c.h) / 2;
            // twiddles e^{-2 pi i k/N} (k < N) then e^{-2 pi i j/(MN)} (j < MN), in double
            const int MN = c.M * N;
            std::vector<float2> tw(N + MN);
            for (int k = 0; k < N; k++) {
                const double ang = -2.0 * M_PI * k / N;
                tw[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
            }
            for (int j = 0; j < MN; j++) {
                const double ang = -2.0 * M_PI * j / MN;
                tw[N + j] = make_float2((float)std::cos(ang), (float)std::sin(ang));
            }
            // frames per batch: spectra of (M + 1) N x N complex per frame within 1 GiB
            const size_t per = (size_t)(c.M + 1) * N * N * sizeof(float2);
            p.chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)c.nF, (1ull << 30) / per));
            p.chunk = std::min(p.chunk, 65535 / (c.M * c.M));   // grid.y = chunk * M^2
            // last pass: row-tiled stores (default) or the pixel-by-pixel kernel; column
            // passes: tiles grouped per XCD (default) or in launch order
            p.scatter = c.var.fht_scatter;
            p.xcd = c.var.fht_xcd;
            p.no_prefetch = c.var.fht_prefetch ? 0 : 1;
            float2 *dtw = nullptr;
            if (!dmalloc(&p.F, sizeof(float2) * (size_t)N * N * p.chunk) ||
                !dmalloc(&p.Y, sizeof(float2) * (size_t)N * N * c.M * p.chunk) ||
                !dmalloc(&dtw, sizeof(float2) * tw.size()))
                return err(ms, ESRRF_ERR_NOMEM, "FHT spectra: hipMalloc failed");
            hipError_t st = hipMemcpy(dtw, tw.data(), sizeof(float2) * tw.size(),
                                      hipMemcpyHostToDevice);
            if (st != hipSuccess) return err(ms, ESRRF_ERR_HIP, "FHT twiddles", st);
            p.tw = dtw;
            p.twMN = dtw + N;
        }
        build_fht_taps(ms, c.radius, c.tSS, c.tSO);
        if (!dmalloc(&ms.d_taps, sizeof(Tap) * ms.h_taps.size()) ||
            !dmalloc(&ms.d_tapk, sizeof(int) * ms.h_tapk.size()))
            return err(ms, ESRRF_ERR_NOMEM, "FHT taps: hipMalloc failed");
        hipError_t st = hipMemcpy(ms.d_taps, ms.h_taps.data(), sizeof(Tap) * ms.h_taps.size(),
                                  hipMemcpyHostToDevice);
        if (st == hipSuccess)
            st = hipMemcpy(ms.d_tapk, ms.h_tapk.data(), sizeof(int) * ms.h_tapk.size(),
                           hipMemcpyHostToDevice);
        if (st != hipSuccess) return err(ms, ESRRF_ERR_HIP, "FHT taps upload", st);
        // gated samples of one frame: sum over taps of the pixels keeping (i, j) in range
        long long act = 0;
        for (int t = 0; t < ms.ntaps; t++) {
            const int i = (ms.h_tapk[t] & 255) - 128, j = ((ms.h_tapk[t] >> 8) & 255) - 128;
            act += (long long)std::max<long>(0, wM - std::abs(i)) * std::max<long>(0, hM - std::abs(j));
        }
        ms.active = act;
        ms.wM = (int)wM;
        ms.hM = (int)hM;
        return ESRRF_OK;
    }
    return mode3d_create(c, ms, d_px, d_U);
}

int mode_prepare(int mode, const ModeCtx &c, ModeState &ms, const void *src, int dtype, int n)
{
    if (mode == 2) return mode3d_prepare(c, ms, src, dtype, n);
    const long npx = (long)c.w * c.h;
    const long nM = (long)c.wM * c.hM;
    hipError_t st = launch_convert_frames(src, dtype, ms.d_rawf, npx * n, c.s);
    if (st != hipSuccess) return err(ms, ESRRF_ERR_HIP, "frame conversion", st);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c.profiling) {
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, c.s);
    }
    if (c.M > 1) {
        st = launch_fht_upsample(ms.plan, ms.d_rawf, c.d_px, n, c.s);
        if (st != hipSuccess) return err(ms, ESRRF_ERR_HIP, "FHT upsampling", st);
        (void)nM;
    } else {
        st = hipMemcpyAsync(c.d_px, ms.d_rawf, sizeof(float) * npx * n, hipMemcpyDeviceToDevice,
                            c.s);
        if (st != hipSuccess) return err(ms, ESRRF_ERR_HIP, "frame copy", st);
    }
    if (c.profiling) {
        (void)hipEventRecord(e1, c.s);
        float ms_ = 0.0f;
        if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms_, e0, e1) == hipSuccess) {
            ms.prof_up_ms += ms_;
            ms.prof_up_n += 1;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    st = launch_fht_prep(c.d_px, c.d_U, c.wM, c.hM, n, c.s);
    if (st != hipSuccess) return err(ms, ESRRF_ERR_HIP, "FHT gradients", st);
    return ESRRF_OK;
}

int mode_pre_rgc(int mode, const ModeCtx &c, ModeState &ms, const ModeRgc &r)
{
    return mode == 2 ? mode3d_lateral(c, ms, r) : ESRRF_OK;
}

int mode_rgc(int mode, const ModeCtx &c, ModeState &ms, const ModeRgc &r)
{
    if (mode == 2) return mode3d_rgc(c, ms, r);
    const int N = r.nFrameForSRRF;
    const float rn = 1.0f / (float)N, rn1 = 1.0f / (float)(N - 1);
    hipError_t st;
    // the tiled kernel keeps the window's tile and the tap table in LDS (64 KiB per block)
    if (r.drift_zero && fht_rgc_lds_bytes(ms.fht_TW, ms.fht_pitch, ms.ntaps) <= 65536) {
        FhtRgcArgs a{};
        a.U = c.d_U;
        a.px = c.d_px;
        a.out = r.out;
        a.prev = r.prev;
        a.taps = ms.d_taps;
        a.tapk = ms.d_tapk;
        a.W = c.wM;
        a.H = c.hM;
        a.frame0 = r.frame0;
        a.nframes = r.nframes;
        a.first_init = r.count0 == 0;
        a.rn = rn;
        a.rn1 = rn1;
        a.sens = r.sens;
        a.sens_int = r.sens_int;
        a.int_weighting = r.int_weighting;
        a.R = ms.fht_R;
        a.Rhi = ms.fht_Rhi;
        a.pair = ms.var.fht_pair;
        a.TW = ms.fht_TW;
        a.TH = ms.fht_TW;
        a.pitch = ms.fht_pitch;
        a.ntaps = ms.ntaps;
        a.wsum = ms.tap_wsum;
        a.tiles_x = (c.wM + 15) / 16;
        a.tiles_y = (c.hM + 15) / 16;
        st = launch_fht_rgc(a, c.s, &ms.rgc_kernel);
    } else {
        FhtGenericArgs g{};
        g.U = c.d_U;
        g.px = c.d_px;
        g.drift = r.drift;
        g.out = r.out;
        g.prev = r.prev;
        g.W = c.wM;
        g.H = c.hM;
        g.frame0 = r.frame0;
        g.nframes = r.nframes;
        g.count0 = r.count0;
        g.nFrameForSRRF = N;
        g.rn = rn;
        g.rn1 = rn1;
        g.sens = r.sens;
        g.sens_int = r.sens_int;
        g.int_weighting = r.int_weighting;
        g.radius = r.radius;
        g.tSS = r.tSS;
        g.tSO = r.tSO;
        st = launch_fht_generic(g, c.s);
        ms.rgc_kernel = "k_fht_generic";
    }
    if (st != hipSuccess) return err(ms, ESRRF_ERR_HIP, "FHT RGC launch", st);
    return ESRRF_OK;
}

int mode_read_gradients(int mode, const ModeCtx &c, ModeState &ms, float *out)
{
    if (mode == 2) return mode3d_read_gradients(c, ms, out);
    const long nM = (long)c.wM * c.hM;
    float *tmp = nullptr;
    if (!dmalloc(&tmp, sizeof(float) * 2 * nM)) return err(ms, ESRRF_ERR_NOMEM, "gradient read");
    hipError_t st = launch_fht_grad(c.d_px, tmp, tmp + nM, c.wM, c.hM, c.s);
    if (st == hipSuccess) st = launch_nan_to_zero_copy(tmp, tmp, 2 * nM, c.s);
    if (st == hipSuccess)
        st = hipMemcpyAsync(out, tmp, sizeof(float) * 2 * nM, hipMemcpyDeviceToHost, c.s);
    if (st == hipSuccess) st = hipStreamSynchronize(c.s);
    (void)hipFree(tmp);
    if (st != hipSuccess) return err(ms, ESRRF_ERR_HIP, "gradient read", st);
    return ESRRF_OK;
}

int mode_reshape_3d(const ModeCtx &c, ModeState &ms, const void *src, int dtype, int n,
                    float *out)
{
    return mode3d_reshape(c, ms, src, dtype, n, out);
}

long long mode_active_samples(int, const ModeState &ms) { return ms.active; }

void free_mode_state(ModeState &ms)
{
    void *ptrs[] = {ms.d_rawf, ms.plan.F,   ms.plan.Y, (void *)ms.plan.tw, ms.d_taps,
                    ms.d_tapk, ms.d_st,     ms.d_cell, ms.d_icf,           ms.d_planes,
                    ms.d_g3,   ms.d_u3,     ms.d_lat,  ms.d_tapv3,        ms.d_tapo3,
                    ms.d_tapk3, ms.d_ph3};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    ms.d_rawf = nullptr;
    ms.plan.F = ms.plan.Y = nullptr;
    ms.plan.tw = nullptr;
    ms.d_taps = nullptr;
    ms.d_tapk = nullptr;
    ms.d_st = nullptr;
    ms.d_cell = nullptr;
    ms.d_icf = nullptr;
    ms.d_planes = ms.d_g3 = ms.d_lat = nullptr;
    ms.d_u3 = nullptr;
    ms.d_tapv3 = nullptr;
    ms.d_tapo3 = ms.d_tapk3 = nullptr;
    ms.d_ph3 = nullptr;
    ms.tiled3 = false;
}

int fht_taps_create(float radius, float tSS, float tSO, FhtTaps &t)
{
    ModeState tmp;
    build_fht_taps(tmp, radius, tSS, tSO);
    t.R = tmp.fht_R;
    t.Rhi = tmp.fht_Rhi;
    t.TW = tmp.fht_TW;
    t.pitch = tmp.fht_pitch;
    t.ntaps = tmp.ntaps;
    t.wsum = tmp.tap_wsum;
    if (!dmalloc(&t.d_taps, sizeof(Tap) * tmp.h_taps.size()) ||
        !dmalloc(&t.d_tapk, sizeof(int) * tmp.h_tapk.size()))
        return ESRRF_ERR_NOMEM;
    hipError_t st = hipMemcpy(t.d_taps, tmp.h_taps.data(), sizeof(Tap) * tmp.h_taps.size(),
                              hipMemcpyHostToDevice);
    if (st == hipSuccess)
        st = hipMemcpy(t.d_tapk, tmp.h_tapk.data(), sizeof(int) * tmp.h_tapk.size(),
                       hipMemcpyHostToDevice);
    return st == hipSuccess ? ESRRF_OK : ESRRF_ERR_HIP;
}

void fht_taps_free(FhtTaps &t)
{
    if (t.d_taps) (void)hipFree(t.d_taps);
    if (t.d_tapk) (void)hipFree(t.d_tapk);
    t.d_taps = nullptr;
    t.d_tapk = nullptr;
}

}  // namespace esrrf
