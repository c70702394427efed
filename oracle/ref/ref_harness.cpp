// ref_harness.cpp -- runs the reference's own OpenCL kernels (compiled from
// /root/reference/resources/*.cl by oracle/ref/build_ref.py into oracle/_ref/*.co) through
// the AMD OpenCL runtime, with the buffer plan, argument binding and enqueue order of
// LiveSRRF_CL (LiveSRRF_CL.java:266-459 BIC/3-D, :462-579 FHT, :636-794).
//
// TEST INFRASTRUCTURE ONLY: what the oracle and the HIP engine are compared against.  Nothing in
// the product links it.  Each .co is specialised for one configuration (the reference
// bakes every constant into the program text, LiveSRRF_CL.java:307-342).
#define CL_TARGET_OPENCL_VERSION 120
#include <CL/cl.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

std::string g_err;

struct Ctx {
    cl_platform_id plat = nullptr;
    cl_device_id dev = nullptr;
    cl_context ctx = nullptr;
    cl_command_queue q = nullptr;
    cl_program prog = nullptr;
    ~Ctx()
    {
        if (prog) clReleaseProgram(prog);
        if (q) clReleaseCommandQueue(q);
        if (ctx) clReleaseContext(ctx);
    }
};

#define CK(x)                                                            \
    do {                                                                 \
        cl_int _e = (x);                                                 \
        if (_e != CL_SUCCESS) {                                          \
            g_err = std::string(#x) + " -> " + std::to_string(_e);       \
            return 1;                                                    \
        }                                                                \
    } while (0)

int open_ctx(Ctx &c, const char *co_path)
{
    cl_uint np = 0;
    CK(clGetPlatformIDs(0, nullptr, &np));
    if (np == 0) {
        g_err = "no OpenCL platform";
        return 1;
    }
    std::vector<cl_platform_id> ps(np);
    CK(clGetPlatformIDs(np, ps.data(), nullptr));
    for (auto p : ps) {
        cl_uint nd = 0;
        if (clGetDeviceIDs(p, CL_DEVICE_TYPE_GPU, 0, nullptr, &nd) == CL_SUCCESS && nd > 0) {
            c.plat = p;
            CK(clGetDeviceIDs(p, CL_DEVICE_TYPE_GPU, 1, &c.dev, nullptr));
            break;
        }
    }
    if (!c.dev) {
        g_err = "no OpenCL GPU device";
        return 1;
    }
    cl_int e;
    c.ctx = clCreateContext(nullptr, 1, &c.dev, nullptr, nullptr, &e);
    CK(e);
    c.q = clCreateCommandQueue(c.ctx, c.dev, 0, &e);
    CK(e);
    FILE *f = fopen(co_path, "rb");
    if (!f) {
        g_err = std::string("cannot open ") + co_path;
        return 1;
    }
    std::vector<unsigned char> bin;
    unsigned char buf[65536];
    size_t r;
    while ((r = fread(buf, 1, sizeof buf, f)) > 0) bin.insert(bin.end(), buf, buf + r);
    fclose(f);
    const unsigned char *bp = bin.data();
    size_t bl = bin.size();
    cl_int bst;
    c.prog = clCreateProgramWithBinary(c.ctx, 1, &c.dev, &bl, &bp, &bst, &e);
    CK(e);
    CK(bst);
    CK(clBuildProgram(c.prog, 1, &c.dev, "", nullptr, nullptr));
    return 0;
}

struct Kern {
    cl_kernel k = nullptr;
    ~Kern()
    {
        if (k) clReleaseKernel(k);
    }
};

struct Buf {
    cl_mem m = nullptr;
    ~Buf()
    {
        if (m) clReleaseMemObject(m);
    }
};

int mk(Ctx &c, Buf &b, size_t bytes)
{
    cl_int e;
    b.m = clCreateBuffer(c.ctx, CL_MEM_READ_WRITE, bytes ? bytes : 4, nullptr, &e);
    CK(e);
    return 0;
}

int kern(Ctx &c, Kern &k, const char *name)
{
    cl_int e;
    k.k = clCreateKernel(c.prog, name, &e);
    if (e != CL_SUCCESS) {
        g_err = std::string("clCreateKernel ") + name + " -> " + std::to_string(e);
        return 1;
    }
    return 0;
}

int args(Kern &k, std::initializer_list<cl_mem> ms)
{
    cl_uint i = 0;
    for (cl_mem m : ms) CK(clSetKernelArg(k.k, i++, sizeof(cl_mem), &m));
    return 0;
}

int run1(Ctx &c, Kern &k, size_t off, size_t n)
{
    if (n == 0) return 0;
    CK(clEnqueueNDRangeKernel(c.q, k.k, 1, &off, &n, nullptr, 0, nullptr, nullptr));
    return 0;
}

int run3(Ctx &c, Kern &k, size_t a, size_t b, size_t d)
{
    size_t off[3] = {0, 0, 0}, g[3] = {a, b, d};
    CK(clEnqueueNDRangeKernel(c.q, k.k, 3, off, g, nullptr, 0, nullptr, nullptr));
    return 0;
}

}  // namespace

extern "C" {

const char *ref_last_error() { return g_err.c_str(); }

// mode 0 = BIC 2-D (liveSRRF.cl), 1 = FHT (liveSRRF_noInterpolation.cl, frames already
// upsampled to wM x hM), 2 = 3-D (live3DSRRF.cl, frames already reshaped to planes).
// loads: per-load frame counts pushed as successive calculateSRRF calls of one SR frame.
// out: [4][nvox] as read by readSRRFbuffer (NaN kept, as in the device buffer);
// mp: [3 * M^d] or NULL; gint: BIC/3-D first-frame GxInt,GyInt(,GzInt) or NULL.
int ref_run(const char *co_path, int mode, int w, int h, int nP, int M, int nFrameForSRRF,
            const float *frames, const int *loads, int nloads, const float *drift,
            const float *st3d, int doMP, int blockLength, float *out, float *mp, float *gint,
            int reset_between_loads)
{
    Ctx c;
    if (open_ctx(c, co_path)) return 1;
    const bool fht = mode == 1, d3 = mode == 2;
    int maxload = 0;
    for (int i = 0; i < nloads; i++) maxload = loads[i] > maxload ? loads[i] : maxload;
    const long wM = (long)w * M, hM = (long)h * M;
    const long single = fht ? wM * hM : (long)w * h * nP;   // pixels of one frame on device
    const long nvox = fht ? wM * hM : wM * hM * (d3 ? (long)nP * M : 1);
    const int gm = 2;
    Buf px, al, gx, gy, gz, gxi, gyi, gzi, prev, outb, dr, cnt, mpb, stb;
    if (mk(c, px, sizeof(float) * single * maxload) || mk(c, gx, sizeof(float) * single * maxload) ||
        mk(c, gy, sizeof(float) * single * maxload) || mk(c, prev, sizeof(float) * nvox) ||
        mk(c, outb, sizeof(float) * 4 * nvox) || mk(c, dr, sizeof(float) * 2 * nFrameForSRRF) ||
        mk(c, cnt, sizeof(int) * 2))
        return 1;
    if (!fht) {
        if (mk(c, gxi, sizeof(float) * gm * gm * single * maxload) ||
            mk(c, gyi, sizeof(float) * gm * gm * single * maxload) ||
            mk(c, mpb, sizeof(float) * 3 * (d3 ? M * M * M : M * M)))
            return 1;
    }
    if (d3) {
        if (mk(c, al, sizeof(float) * single * maxload) || mk(c, gz, sizeof(float) * single * maxload) ||
            mk(c, gzi, sizeof(float) * gm * gm * single * maxload) || mk(c, stb, sizeof(float) * 3 * nP))
            return 1;
        CK(clEnqueueWriteBuffer(c.q, stb.m, CL_TRUE, 0, sizeof(float) * 3 * nP, st3d, 0, nullptr, nullptr));
    }
    Kern kA, kG, kI, kR, kInc, kRst, kVar, kMP, kCor;
    if (kern(c, kG, fht ? "calculateGradientRobX" : "calculateGradient_2point") ||
        kern(c, kR, "calculateRadialGradientConvergence") ||
        kern(c, kInc, "kernelIncrementFramePosition") || kern(c, kRst, "kernelResetFramePosition") ||
        kern(c, kVar, "kernelCalculateVar"))
        return 1;
    if (!fht && (kern(c, kI, "calculateGradientInterpolation") || kern(c, kMP, "kernelCalculateMPmap") ||
                 kern(c, kCor, "kernelCorrectMPmap")))
        return 1;
    if (d3 && kern(c, kA, "kernelAlignPixels")) return 1;
    // argument binding of LiveSRRF_CL.java:366-418 / :544-564
    if (d3) {
        if (args(kA, {px.m, al.m, stb.m}) || args(kG, {al.m, gx.m, gy.m, gz.m, cnt.m}) ||
            args(kI, {gx.m, gy.m, gz.m, gxi.m, gyi.m, gzi.m}) ||
            args(kR, {al.m, gxi.m, gyi.m, gzi.m, prev.m, outb.m, dr.m, cnt.m}))
            return 1;
    } else if (fht) {
        if (args(kG, {px.m, gx.m, gy.m, cnt.m}) ||
            args(kR, {px.m, gx.m, gy.m, prev.m, outb.m, dr.m, cnt.m}))
            return 1;
    } else {
        if (args(kG, {px.m, gx.m, gy.m, cnt.m}) || args(kI, {gx.m, gy.m, gxi.m, gyi.m}) ||
            args(kR, {px.m, gxi.m, gyi.m, prev.m, outb.m, dr.m, cnt.m}))
            return 1;
    }
    if (args(kInc, {cnt.m}) || args(kRst, {cnt.m}) || args(kVar, {outb.m})) return 1;
    if (!fht && (args(kMP, {outb.m, mpb.m}) || args(kCor, {outb.m, mpb.m}))) return 1;

    // fresh buffers: zero them (OpenCL leaves them undefined); drift x M already applied
    std::vector<float> zero(4 * nvox, 0.0f);
    CK(clEnqueueWriteBuffer(c.q, outb.m, CL_TRUE, 0, sizeof(float) * 4 * nvox, zero.data(), 0, nullptr, nullptr));
    CK(clEnqueueWriteBuffer(c.q, prev.m, CL_TRUE, 0, sizeof(float) * nvox, zero.data(), 0, nullptr, nullptr));
    int zc[2] = {0, 0};
    CK(clEnqueueWriteBuffer(c.q, cnt.m, CL_TRUE, 0, sizeof zc, zc, 0, nullptr, nullptr));
    CK(clEnqueueWriteBuffer(c.q, dr.m, CL_TRUE, 0, sizeof(float) * 2 * nFrameForSRRF, drift, 0, nullptr, nullptr));
    // resetFramePosition (:955-960)
    if (run1(c, kRst, 0, 1)) return 1;

    const long rgcN = fht ? wM * hM : (d3 ? (long)nP * M * wM * hM : wM * hM);
    const long nBlocks = rgcN / blockLength + ((rgcN % blockLength) == 0 ? 0 : 1);
    const float *src = frames;
    for (int l = 0; l < nloads; l++) {
        const int n = loads[l];
        if (reset_between_loads && l > 0 && run1(c, kRst, 0, 1)) return 1;
        CK(clEnqueueWriteBuffer(c.q, px.m, CL_TRUE, 0, sizeof(float) * single * n, src, 0, nullptr, nullptr));
        src += single * n;
        if (d3) {
            CK(clFinish(c.q));
            if (run1(c, kA, 0, (size_t)single * n)) return 1;
            CK(clFinish(c.q));
            if (run1(c, kG, 0, (size_t)single * n)) return 1;
            CK(clFinish(c.q));
            if (run1(c, kI, 0, (size_t)gm * gm * single * n)) return 1;
        } else if (fht) {
            if (run3(c, kG, wM, hM, n)) return 1;
        } else {
            CK(clFinish(c.q));
            if (run3(c, kG, w, h, n)) return 1;
            CK(clFinish(c.q));
            if (run3(c, kI, (size_t)gm * w, (size_t)gm * h, n)) return 1;
        }
        CK(clFinish(c.q));
        if (gint && l == 0 && !fht) {
            const long gs = (long)gm * gm * w * h * nP;   // first frame, every plane
            CK(clEnqueueReadBuffer(c.q, gxi.m, CL_TRUE, 0, sizeof(float) * gs, gint, 0, nullptr, nullptr));
            CK(clEnqueueReadBuffer(c.q, gyi.m, CL_TRUE, 0, sizeof(float) * gs, gint + gs, 0, nullptr, nullptr));
            if (d3)
                CK(clEnqueueReadBuffer(c.q, gzi.m, CL_TRUE, 0, sizeof(float) * gs, gint + 2 * gs, 0, nullptr, nullptr));
        }
        // per frame: RGC in blocks of blockLength with a global offset, then increment
        // (LiveSRRF_CL.java:713-742)
        for (int f = 0; f < n; f++) {
            for (long b = 0; b < nBlocks; b++) {
                long ws = rgcN - b * blockLength;
                if (ws > blockLength) ws = blockLength;
                if (run1(c, kR, (size_t)(b * blockLength), (size_t)ws)) return 1;
            }
            CK(clFinish(c.q));
            if (run1(c, kInc, 0, 2)) return 1;
        }
    }
    // readSRRFbuffer (:749-794)
    CK(clFinish(c.q));
    const long nDim = d3 ? (long)M * M * M : (long)M * M;
    if (run1(c, kVar, 0, (size_t)(fht ? wM * hM : nDim * single))) return 1;
    if (doMP && !fht) {
        if (run1(c, kMP, 0, (size_t)(3 * nDim))) return 1;
        if (run1(c, kCor, 0, (size_t)(3 * nDim * single))) return 1;
    }
    CK(clFinish(c.q));
    CK(clEnqueueReadBuffer(c.q, outb.m, CL_TRUE, 0, sizeof(float) * 4 * nvox, out, 0, nullptr, nullptr));
    if (mp && doMP && !fht)
        CK(clEnqueueReadBuffer(c.q, mpb.m, CL_TRUE, 0, sizeof(float) * 3 * nDim, mp, 0, nullptr, nullptr));
    return 0;
}

}  // extern "C"
