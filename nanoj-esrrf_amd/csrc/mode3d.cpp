// mode3d.cpp -- host side of the 3-D (multi-focus) pipeline: calibration tables
// (LiveSRRF_CL.java:199-245, :438-458), buffers, and the per-load / per-launch sequence of
// calculateSRRF's 3-D branch (:655-696, :699-746).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "geometry.h"
#include "mode3d_internal.h"

namespace esrrf {

namespace {

int err3(ModeState &ms, int code, const char *msg, hipError_t st = hipSuccess)
{
    ms.err = msg;
    if (st != hipSuccess) {
        ms.err += ": ";
        ms.err += hipGetErrorString(st);
    }
    return code;
}

template <class T>
bool alloc(T **p, size_t bytes)
{
    *p = nullptr;
    if (hipMalloc((void **)p, bytes ? bytes : 4) == hipSuccess) return true;
    (void)hipGetLastError();   // a failed allocation must not surface in later checks
    return false;
}

Prep3dArgs prep_args(const ModeCtx &c, const ModeState &ms)
{
    Prep3dArgs p{};
    p.W = c.W;
    p.H = c.H;
    p.w = c.w;
    p.h = c.h;
    p.nP = c.nP;
    p.cell = ms.d_cell;
    p.ic = ms.d_icf;
    p.bg = ms.bg;
    p.st = ms.d_st;
    p.planes = ms.d_planes;
    p.aligned = c.d_px;
    p.g = ms.d_g3;
    p.U = ms.d_u3;
    return p;
}

// Tap tables of the tiled RGC kernel (frames without drift).  Where the window, the tile
// or the table size exceed the kernel's limits the per-voxel kernels serve every frame.
int build_tiled3(const ModeCreate &c, ModeState &ms)
{
    ms.tiled3 = false;
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return err3(ms, ESRRF_ERR_HIP, "3-D: device LDS limit query failed");
    ms.lds_limit = (size_t)lds;
    Geometry3 g;
    if (!build_3d_geometry(c.w, c.h, c.nP, c.M, c.Mz, c.radius, c.tSS, c.tSO, g)) return ESRRF_OK;
    const size_t ntap = g.tapv.size();
    if (ntap > ((size_t)1 << 28) || g.NZ * g.TH * g.TW > 256 * 16) return ESRRF_OK;
    if (!alloc(&ms.d_tapv3, sizeof(float4) * ntap) || !alloc(&ms.d_tapo3, sizeof(int) * ntap) ||
        !alloc(&ms.d_tapk3, sizeof(int) * ntap) ||
        !alloc(&ms.d_ph3, sizeof(Phase3) * g.phases.size()))
        return err3(ms, ESRRF_ERR_NOMEM, "3-D tap tables: hipMalloc failed");
    hipError_t e = hipMemcpy(ms.d_tapv3, g.tapv.data(), sizeof(float4) * ntap,
                             hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(ms.d_tapo3, g.tapo.data(), sizeof(int) * ntap, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(ms.d_tapk3, g.tapk.data(), sizeof(int) * ntap, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(ms.d_ph3, g.phases.data(), sizeof(Phase3) * g.phases.size(),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) return err3(ms, ESRRF_ERR_HIP, "3-D tap tables upload", e);
    Rgc3Tables &t = ms.t3;
    t.tapv = ms.d_tapv3;
    t.tapo = ms.d_tapo3;
    t.tapk = ms.d_tapk3;
    t.phases = ms.d_ph3;
    t.R2lo = g.R2lo;
    t.R2hi = g.R2hi;
    t.G0 = g.G0;
    t.TW = g.TW;
    t.TH = g.TH;
    t.PW = g.PW;
    t.P = g.P;
    t.PZ = g.PZ;
    t.NZ = g.NZ;
    t.xs_log2 = g.xs_log2;
    t.ncls_x = g.ncls_x;
    t.ncls_y = g.ncls_y;
    t.nZM = g.nZM;
    t.gtaps = g.max_group_taps + TAP_SPARE;
    // two z-slices per block when their union tile fits the staging plan (4096 elements)
    t.NZ2 = g.pair_ok && g.NZ2 * g.TH * g.TW <= 256 * 16 ? g.NZ2 : 0;
    ms.tiled3 = true;
    return ESRRF_OK;
}

}  // namespace

int mode3d_create(const ModeCreate &c, ModeState &ms, float **d_px, float2 **d_U)
{
    const esrrf_params *p = c.p;
    const int nP = c.nP, w = c.w, h = c.h, M = c.M, nF = c.nF;
    *d_px = nullptr;
    *d_U = nullptr;
    if (nP < 2)
        return err3(ms, ESRRF_ERR_UNSUPPORTED,
                    "3-D mode needs at least 2 planes (the axial cubic reads planes z0..z0+1)");
    // shiftXYtheta3D (LiveSRRF_CL.java:440-445)
    std::vector<float> st(3 * nP);
    for (int i = 0; i < 3 * nP; i++) {
        if (i / nP == 0) st[i] = (float)(-p->theta_3d[i] * M_PI / 180.0);
        if (i / nP == 1) st[i] = (float)p->shift_x_3d[i - nP];
        if (i / nP == 2) st[i] = (float)p->shift_y_3d[i - 2 * nP];
    }
    // montage cells and intensity scaling (LiveSRRF_CL.java:972-986)
    std::vector<int> cell(2 * nP);
    std::vector<float> icf(nP);
    for (int z = 0; z < nP; z++) {
        const int roi = (int)p->roi_3d[z];
        const int i = roi % 3;
        const int j = (int)(p->roi_3d[z] - i) / 3;
        const int x = (int)std::floor((float)c.W / 3.0f * (float)i + 0.5f);
        const int y = (int)std::floor((float)c.H / 3.0f * (float)j + 0.5f);
        if (i < 0 || j < 0 || x + w > c.W || y + h > c.H)
            return err3(ms, ESRRF_ERR_INVALID, "3-D calibration: ROI # outside the 3x3 montage");
        cell[z] = x;
        cell[nP + z] = y;
        icf[z] = (float)p->int_coeffs_3d[z];
    }
    ms.bg = p->background_3d;
    const long wh = (long)w * h;
    const long npl = (long)nF * nP;
    const long whM = (long)w * M * h * M;
    if (!alloc(d_px, sizeof(float) * wh * npl) || !alloc(&ms.d_planes, sizeof(float) * wh * npl) ||
        !alloc(&ms.d_g3, sizeof(float) * 3 * wh * npl) ||
        !alloc(&ms.d_u3, sizeof(float4) * 4 * wh * npl) ||
        !alloc(&ms.d_lat, sizeof(float) * whM * npl) || !alloc(&ms.d_st, sizeof(float) * 3 * nP) ||
        !alloc(&ms.d_cell, sizeof(int) * 2 * nP) || !alloc(&ms.d_icf, sizeof(float) * nP))
        return err3(ms, ESRRF_ERR_NOMEM, "3-D buffers: hipMalloc failed");
    hipError_t e = hipMemcpy(ms.d_st, st.data(), sizeof(float) * 3 * nP, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(ms.d_cell, cell.data(), sizeof(int) * 2 * nP, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(ms.d_icf, icf.data(), sizeof(float) * nP, hipMemcpyHostToDevice);
    if (e != hipSuccess) return err3(ms, ESRRF_ERR_HIP, "3-D tables upload", e);
    // gated samples of one frame without drift (geometry only; used for FLOP accounting)
    unsigned long long *d_cnt = nullptr;
    if (!alloc(&d_cnt, sizeof(unsigned long long)))
        return err3(ms, ESRRF_ERR_NOMEM, "3-D sample count: hipMalloc failed");
    unsigned long long cnt = 0;
    e = hipMemcpy(d_cnt, &cnt, sizeof(cnt), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch3d_count(w, h, nP, M, c.Mz, c.radius, c.tSO, d_cnt, 0);
    if (e == hipSuccess) e = hipMemcpy(&cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost);
    (void)hipFree(d_cnt);
    if (e != hipSuccess) return err3(ms, ESRRF_ERR_HIP, "3-D sample count", e);
    ms.active = (long long)cnt;
    ms.Mz = c.Mz;
    ms.wM = w * M;
    ms.hM = h * M;
    return build_tiled3(c, ms);
}

int mode3d_prepare(const ModeCtx &c, ModeState &ms, const void *src, int dtype, int n)
{
    const Prep3dArgs p = prep_args(c, ms);
    hipError_t e = launch3d_prepare(src, dtype, p, n, c.s);
    if (e != hipSuccess) return err3(ms, ESRRF_ERR_HIP, "3-D reshape/align/gradients", e);
    ms.n_loaded = n;
    return ESRRF_OK;
}

// The planes buffer is the reshape stage of every load; the SR state (aligned planes,
// gradients, cumulants) does not depend on its contents after the load, so it serves as
// the device side of this read-back.
int mode3d_reshape(const ModeCtx &c, ModeState &ms, const void *src, int dtype, int n,
                   float *out)
{
    const Prep3dArgs p = prep_args(c, ms);
    hipError_t e = launch3d_reshape(src, dtype, p, n, c.s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, ms.d_planes, sizeof(float) * (size_t)n * c.nP * c.w * c.h,
                           hipMemcpyDeviceToHost, c.s);
    if (e == hipSuccess) e = hipStreamSynchronize(c.s);
    if (e != hipSuccess) return err3(ms, ESRRF_ERR_HIP, "3-D reshape", e);
    return ESRRF_OK;
}

static Rgc3dArgs rgc3_args(const ModeCtx &c, ModeState &ms, const ModeRgc &r)
{
    Rgc3dArgs a{};
    a.U = ms.d_u3;
    a.px = c.d_px;
    a.L = ms.d_lat;
    a.drift = r.drift;
    a.out = r.out;
    a.prev = r.prev;
    a.w = c.w;
    a.h = c.h;
    a.nP = c.nP;
    a.M = c.M;
    a.Mz = ms.Mz;
    a.wM = c.wM;
    a.hM = c.hM;
    a.frame0 = r.frame0;
    a.nframes = r.nframes;
    a.count0 = r.count0;
    a.nFor = r.nFrameForSRRF;
    a.sens = r.sens;
    a.sens_int = r.sens_int;
    a.int_weighting = r.int_weighting;
    a.radius = r.radius;
    a.tSS = r.tSS;
    a.tSO = r.tSO;
    return a;
}

int mode3d_lateral(const ModeCtx &c, ModeState &ms, const ModeRgc &r)
{
    hipError_t e = launch3d_lateral(rgc3_args(c, ms, r), c.s);
    if (e != hipSuccess) return err3(ms, ESRRF_ERR_HIP, "3-D lateral interpolation", e);
    return ESRRF_OK;
}

int mode3d_rgc(const ModeCtx &c, ModeState &ms, const ModeRgc &r)
{
    const Rgc3dArgs a = rgc3_args(c, ms, r);
    hipError_t e = launch3d_rgc(a, r.drift_zero, ms.tiled3 ? &ms.t3 : nullptr, ms.lds_limit,
                                c.s, &ms.rgc_kernel, ms.var.k3d_kernel, ms.var.k3d_split,
                                ms.var.k3d_zpair);
    if (e != hipSuccess) return err3(ms, ESRRF_ERR_HIP, "3-D RGC launch", e);
    return ESRRF_OK;
}

int mode3d_read_gradients(const ModeCtx &c, ModeState &ms, float *out)
{
    const long n = (long)c.nP * 3 * 4 * c.w * c.h;
    float *tmp = nullptr;
    if (!alloc(&tmp, sizeof(float) * n)) return err3(ms, ESRRF_ERR_NOMEM, "gradient read");
    // the gradient planes of a load are stored [3][n_loaded * nP][h][w]
    const long gn = (long)ms.n_loaded * c.nP * c.w * c.h;
    hipError_t e = launch3d_gint(ms.d_g3, tmp, c.w, c.h, c.nP, gn, c.s);
    if (e == hipSuccess) e = launch_nan_to_zero_copy(tmp, tmp, n, c.s);
    if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, sizeof(float) * n, hipMemcpyDeviceToHost, c.s);
    if (e == hipSuccess) e = hipStreamSynchronize(c.s);
    (void)hipFree(tmp);
    if (e != hipSuccess) return err3(ms, ESRRF_ERR_HIP, "gradient read", e);
    return ESRRF_OK;
}

}  // namespace esrrf
