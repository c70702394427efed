// device_interp.h -- the reference's interpolation functions as device code, shared by
// the BIC, FHT and 3-D kernels (liveSRRF.cl:26-133; identical copies in
// liveSRRF_noInterpolation.cl:17-75 and live3DSRRF.cl:30-140).
#pragma once
#include <hip/hip_runtime.h>

namespace esrrf {

// Catmull-Rom kernel, same expression order as liveSRRF.cl:26-35 (exact for the half-grid
// arguments the gradient interpolation uses).
__device__ __forceinline__ float cubic_w(float x)
{
    const float a = 0.5f;
    if (x < 0.0f) x = -x;
    float z = 0.0f;
    if (x < 1.0f) {
        float t = __fmul_rn(x, -a + 2.0f);
        t = __fadd_rn(t, a - 3.0f);
        z = __fmul_rn(x, x);
        z = __fmul_rn(z, t);
        z = __fadd_rn(z, 1.0f);
    } else if (x < 2.0f) {
        float t1 = __fmul_rn(__fmul_rn(__fmul_rn(-a, x), x), x);
        float t2 = __fmul_rn(__fmul_rn(5.0f * a, x), x);
        float t3 = __fmul_rn(8.0f * a, x);
        z = __fadd_rn(__fsub_rn(__fadd_rn(t1, t2), t3), 4.0f * a);
    }
    return z;
}

// getInterpolatedValue on a plane stored in a (possibly offset) tile:
// value of image pixel (u, v) is tile[(v - oy) * pitch + (u - ox)].
__device__ __forceinline__ float interp_tile(const float *t, int pitch, int ox, int oy, int w,
                                             int h, float x, float y)
{
    const int u0 = (int)floorf(x);
    const int v0 = (int)floorf(y);
    float q = 0.0f;
    if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2) {
#pragma unroll
        for (int j = 0; j <= 3; j++) {
            const int v = v0 - 1 + j;
            float p = 0.0f;
#pragma unroll
            for (int i = 0; i <= 3; i++) {
                const int u = u0 - 1 + i;
                p = fmaf(t[(v - oy) * pitch + (u - ox)], cubic_w(x - (float)u), p);
            }
            q = fmaf(p, cubic_w(y - (float)v), q);
        }
    } else {
        const int xbase = (int)fminf((float)w - 2.0f, fmaxf(x, 0.0f));
        const int ybase = (int)fminf((float)h - 2.0f, fmaxf(y, 0.0f));
        const float xf = x - (float)xbase;
        const float yf = y - (float)ybase;
        const float ll = t[(ybase - oy) * pitch + (xbase - ox)];
        const float lr = t[(ybase - oy) * pitch + (xbase + 1 - ox)];
        const float ur = t[(ybase + 1 - oy) * pitch + (xbase + 1 - ox)];
        const float ul = t[(ybase + 1 - oy) * pitch + (xbase - ox)];
        const float ua = fmaf(xf, ur - ul, ul);
        const float la = fmaf(xf, lr - ll, ll);
        q = fmaf(yf, ua - la, la);
    }
    return q;
}

// interp_tile on a plane whose pixels are every S-th float of the tile (one plane of S
// interleaved planes), same operations.
template <int S>
__device__ __forceinline__ float interp_tile_s(const float *t, int pitch, int ox, int oy, int w,
                                               int h, float x, float y)
{
    const int u0 = (int)floorf(x);
    const int v0 = (int)floorf(y);
    float q = 0.0f;
    if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2) {
#pragma unroll
        for (int j = 0; j <= 3; j++) {
            const int v = v0 - 1 + j;
            float p = 0.0f;
#pragma unroll
            for (int i = 0; i <= 3; i++) {
                const int u = u0 - 1 + i;
                p = fmaf(t[S * ((v - oy) * pitch + (u - ox))], cubic_w(x - (float)u), p);
            }
            q = fmaf(p, cubic_w(y - (float)v), q);
        }
    } else {
        const int xbase = (int)fminf((float)w - 2.0f, fmaxf(x, 0.0f));
        const int ybase = (int)fminf((float)h - 2.0f, fmaxf(y, 0.0f));
        const float xf = x - (float)xbase;
        const float yf = y - (float)ybase;
        const float ll = t[S * ((ybase - oy) * pitch + (xbase - ox))];
        const float lr = t[S * ((ybase - oy) * pitch + (xbase + 1 - ox))];
        const float ur = t[S * ((ybase + 1 - oy) * pitch + (xbase + 1 - ox))];
        const float ul = t[S * ((ybase + 1 - oy) * pitch + (xbase - ox))];
        const float ua = fmaf(xf, ur - ul, ul);
        const float la = fmaf(xf, lr - ll, ll);
        q = fmaf(yf, ua - la, la);
    }
    return q;
}

__device__ __forceinline__ float interp_global(const float *__restrict__ a, int w, int h,
                                               float x, float y)
{
    const int u0 = (int)floorf(x);
    const int v0 = (int)floorf(y);
    float q = 0.0f;
    if (u0 > 0 && u0 < w - 2 && v0 > 0 && v0 < h - 2) {
        for (int j = 0; j <= 3; j++) {
            const int v = v0 - 1 + j;
            float p = 0.0f;
            for (int i = 0; i <= 3; i++) {
                const int u = u0 - 1 + i;
                p = p + a[(long)v * w + u] * cubic_w(x - (float)u);
            }
            q = q + p * cubic_w(y - (float)v);
        }
    } else {
        const int xbase = (int)fminf((float)w - 2.0f, fmaxf(x, 0.0f));
        const int ybase = (int)fminf((float)h - 2.0f, fmaxf(y, 0.0f));
        const float xf = x - (float)xbase;
        const float yf = y - (float)ybase;
        const float ll = a[(long)ybase * w + xbase];
        const float lr = a[(long)ybase * w + xbase + 1];
        const float ur = a[(long)(ybase + 1) * w + xbase + 1];
        const float ul = a[(long)(ybase + 1) * w + xbase];
        const float ua = ul + xf * (ur - ul);
        const float la = ll + xf * (lr - ll);
        q = la + yf * (ua - la);
    }
    return q;
}


}  // namespace esrrf
