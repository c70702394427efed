// k_rse.hip -- SQUIRREL resolution-scaled error of the AVG image and its gradient with
// respect to (fwhm, sensitivity, RSF sigma) (ErrorMapLiveSRRF.java:81-246).
//
// Inputs are the four channels of the adjoint RGC launch: AVG, dAVG/dtSS, dAVG/dS, INT.
//   k_rse_mpstats  per sub-pixel position: means of AVG, dAVG/dtSS, dAVG/dS over all
//                  macro-pixels (the MP map of liveSRRF.cl:464-485, with its derivatives)
//   k_rse_prep     crop; MP-corrected SR = AVG/m and its forward-mode derivatives
//                  d(AVG/m) = dAVG/m - AVG dm/m^2; ref = INT
//   k_gauss_taps   normalised Gaussian taps and their sigma-derivative (one block)
//   k_blur_cols / k_blur_rows   separable blur, clamp-to-edge
//   k_rse_sums     sums for the least-squares alpha, beta (ErrorMapLiveSRRF.java:201-233)
//   k_rse_resid    residual r = ref - alpha x - beta: sum r^2 and the gradient dot products
// Every sum is in double; partial sums per block, then one block finishes them, so a value
// and gradient cost one read-back of a few doubles.
#include "rse_internal.h"

namespace esrrf {

constexpr int RB = 256;

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = RB / 2; s > 0; s >>= 1) {
        if (tid < s) sh[tid] += sh[tid + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// one block per sub-pixel position (sy, sx)
__global__ void __launch_bounds__(RB) k_rse_mpstats(const float *__restrict__ out4, int w, int h,
                                                    int M, double *__restrict__ stats)
{
    __shared__ double sh[RB];
    const int s = blockIdx.x;
    const int sy = s / M, sx = s - sy * M;
    const long whM = (long)w * M * h * M;
    const int wM = w * M;
    double a = 0.0, b = 0.0, c = 0.0;
    for (long i = threadIdx.x; i < (long)w * h; i += RB) {
        const int y = (int)(i / w), x = (int)(i - (long)y * w);
        const long p = (long)(y * M + sy) * wM + (x * M + sx);
        a += out4[p];
        b += out4[p + whM];
        c += out4[p + 2 * whM];
    }
    const double n = (double)w * h;
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    c = block_sum(c, sh);
    if (threadIdx.x == 0) {
        stats[s] = a / n;
        stats[M * M + s] = b / n;
        stats[2 * M * M + s] = c / n;
    }
}

__global__ void k_rse_prep(const float *__restrict__ out4, const double *__restrict__ stats,
                           int w, int h, int M, int crop, int mp, float *__restrict__ Y,
                           float *__restrict__ SR, float *__restrict__ Dt, float *__restrict__ DS)
{
    const int wM = w * M, hM = h * M;
    const int Wc = wM - 2 * crop, Hc = hM - 2 * crop;
    const long n = (long)Wc * Hc;
    const long whM = (long)wM * hM;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const int yc = (int)(i / Wc), xc = (int)(i - (long)yc * Wc);
        const int yM = yc + crop, xM = xc + crop;
        const long p = (long)yM * wM + xM;
        const float avg = out4[p], gt = out4[p + whM], gs = out4[p + 2 * whM];
        Y[i] = out4[p + 3 * whM];
        if (mp) {
            const int s = (yM % M) * M + (xM % M);
            const float m = (float)stats[s];
            const float mt = (float)stats[M * M + s], ms = (float)stats[2 * M * M + s];
            const float sr = avg / m;
            if (isnan(sr)) {
                // a sub-pixel position without gated samples has m = 0: readSRRFbuffer
                // hands the 0/0 over as 0 (LiveSRRF_CL.java:781), constant in the parameters
                SR[i] = 0.0f;
                Dt[i] = 0.0f;
                DS[i] = 0.0f;
            } else {
                SR[i] = sr;
                Dt[i] = (gt - sr * mt) / m;
                DS[i] = (gs - sr * ms) / m;
            }
        } else {
            SR[i] = avg;
            Dt[i] = gt;
            DS[i] = gs;
        }
    }
}

// taps[0..2K] = g, taps[2K+1..4K+1] = dg/dsigma
__global__ void __launch_bounds__(RB) k_gauss_taps(float sigma, int K, float *__restrict__ taps)
{
    __shared__ double sh[RB];
    const int n = 2 * K + 1;
    double ez = 0.0, dz = 0.0;
    const double s = (double)sigma;
    for (int i = threadIdx.x; i < n; i += RB) {
        const double k = (double)(i - K);
        const double e = s > 0.0 ? exp(-k * k / (2.0 * s * s)) : (i == K ? 1.0 : 0.0);
        ez += e;
        dz += s > 0.0 ? e * k * k / (s * s * s) : 0.0;
    }
    const double Z = block_sum(ez, sh);
    const double dZ = block_sum(dz, sh);
    for (int i = threadIdx.x; i < n; i += RB) {
        const double k = (double)(i - K);
        const double e = s > 0.0 ? exp(-k * k / (2.0 * s * s)) : (i == K ? 1.0 : 0.0);
        const double de = s > 0.0 ? e * k * k / (s * s * s) : 0.0;
        taps[i] = (float)(e / Z);
        taps[n + i] = (float)((de * Z - e * dZ) / (Z * Z));
    }
}

// The value taps of k_gauss_taps, computed by one block into LDS with the same arithmetic
// and reduction (so the same bits); for the blur kernels of the value-only evaluation.
__device__ void gauss_taps_block(float sigma, int K, float *t, double *sh)
{
    const int n = 2 * K + 1;
    double ez = 0.0;
    const double s = (double)sigma;
    for (int i = threadIdx.x; i < n; i += RB) {
        const double k = (double)(i - K);
        ez += s > 0.0 ? exp(-k * k / (2.0 * s * s)) : (i == K ? 1.0 : 0.0);
    }
    const double Z = block_sum(ez, sh);
    for (int i = threadIdx.x; i < n; i += RB) {
        const double k = (double)(i - K);
        const double e = s > 0.0 ? exp(-k * k / (2.0 * s * s)) : (i == K ? 1.0 : 0.0);
        t[i] = (float)(e / Z);
    }
    __syncthreads();
}

// k_blur_cols / k_blur_rows (one image) with the taps of sigma formed in LDS
__global__ void __launch_bounds__(RB) k_blur_cols_sig(const float *__restrict__ in,
                                                      float *__restrict__ out, int W, int H,
                                                      float sigma, int K)
{
    extern __shared__ float tl[];
    __shared__ double sh[RB];
    gauss_taps_block(sigma, K, tl, sh);
    const long n = (long)W * H;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const int y = (int)(i / W), x = (int)(i - (long)y * W);
        float acc = 0.0f;
        for (int k = -K; k <= K; k++) {
            const int yy = min(max(y + k, 0), H - 1);
            acc = fmaf(tl[k + K], in[(long)yy * W + x], acc);
        }
        out[i] = acc;
    }
}

__global__ void __launch_bounds__(RB) k_blur_rows_sig(const float *__restrict__ a,
                                                      float *__restrict__ out, int W, int H,
                                                      float sigma, int K)
{
    extern __shared__ float tl[];
    __shared__ double sh[RB];
    gauss_taps_block(sigma, K, tl, sh);
    const long n = (long)W * H;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const int y = (int)(i / W), x = (int)(i - (long)y * W);
        const float *ra = a + (long)y * W;
        float acc = 0.0f;
        for (int k = -K; k <= K; k++) acc = fmaf(tl[k + K], ra[min(max(x + k, 0), W - 1)], acc);
        out[i] = acc;
    }
}

// out = sum_k taps[k] in[clamp(y + k - K)][x] (taps staged in LDS)
__global__ void __launch_bounds__(RB) k_blur_cols(const float *__restrict__ in,
                                                  float *__restrict__ out, int W, int H,
                                                  const float *__restrict__ taps_g, int K)
{
    extern __shared__ float taps[];
    for (int i = threadIdx.x; i < 2 * K + 1; i += blockDim.x) taps[i] = taps_g[i];
    __syncthreads();
    const long n = (long)W * H;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const int y = (int)(i / W), x = (int)(i - (long)y * W);
        float acc = 0.0f;
        for (int k = -K; k <= K; k++) {
            const int yy = min(max(y + k, 0), H - 1);
            acc = fmaf(taps[k + K], in[(long)yy * W + x], acc);
        }
        out[i] = acc;
    }
}

// out = sum_k ta[k] a[y][clamp(x + k - K)] (+ sum_k tb[k] b[...] when b is given; taps
// staged in LDS)
__global__ void __launch_bounds__(RB) k_blur_rows(const float *__restrict__ a,
                                                  const float *__restrict__ ta_g,
                                                  const float *__restrict__ b,
                                                  const float *__restrict__ tb_g,
                                                  float *__restrict__ out, int W, int H, int K)
{
    extern __shared__ float tl[];
    float *ta = tl, *tb = tl + 2 * K + 1;
    for (int i = threadIdx.x; i < 2 * K + 1; i += blockDim.x) {
        ta[i] = ta_g[i];
        if (b) tb[i] = tb_g[i];
    }
    __syncthreads();
    const long n = (long)W * H;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const int y = (int)(i / W), x = (int)(i - (long)y * W);
        const float *ra = a + (long)y * W;
        float acc = 0.0f;
        for (int k = -K; k <= K; k++) acc = fmaf(ta[k + K], ra[min(max(x + k, 0), W - 1)], acc);
        if (b) {
            const float *rb = b + (long)y * W;
            for (int k = -K; k <= K; k++)
                acc = fmaf(tb[k + K], rb[min(max(x + k, 0), W - 1)], acc);
        }
        out[i] = acc;
    }
}

// partial sums per block: sum y, sum x, sum x^2, sum x y
__global__ void __launch_bounds__(RB) k_rse_sums(const float *__restrict__ Y,
                                                 const float *__restrict__ X, long n,
                                                 double *__restrict__ part)
{
    __shared__ double sh[RB];
    double sy = 0, sx = 0, sxx = 0, sxy = 0;
    for (long i = (long)blockIdx.x * RB + threadIdx.x; i < n; i += (long)gridDim.x * RB) {
        const double y = Y[i], x = X[i];
        sy += y;
        sx += x;
        sxx += x * x;
        sxy += x * y;
    }
    sy = block_sum(sy, sh);
    sx = block_sum(sx, sh);
    sxx = block_sum(sxx, sh);
    sxy = block_sum(sxy, sh);
    if (threadIdx.x == 0) {
        double *p = part + 4 * blockIdx.x;
        p[0] = sy;
        p[1] = sx;
        p[2] = sxx;
        p[3] = sxy;
    }
}

// red[10] = sum y^2 of the cropped reference (once per reference image), from the sums
// pass run with X = Y (its third partial)
__global__ void __launch_bounds__(RB) k_rse_yy(const double *__restrict__ part, int nparts,
                                               double *__restrict__ red)
{
    __shared__ double sh[RB];
    double t = 0;
    for (int b = threadIdx.x; b < nparts; b += RB) t += part[4 * b + 2];
    t = block_sum(t, sh);
    if (threadIdx.x == 0) red[10] = t;
}

// red[0..3] = totals; red[4] = alpha, red[5] = beta (ErrorMapLiveSRRF.java:201-233 with the
// blurred ones equal to 1, so N = sum 1^2 = n).  closed: also red[6] = sum of squared
// residuals = sum y^2 - alpha sum xy - beta sum y, exact at the least-squares alpha, beta
// (the value-only evaluation needs no residual pass)
__global__ void __launch_bounds__(RB) k_rse_alpha_beta(const double *__restrict__ part,
                                                       int nparts, long n,
                                                       double *__restrict__ red, int closed)
{
    __shared__ double sh[RB];
    double t[4] = {0, 0, 0, 0};
    for (int b = threadIdx.x; b < nparts; b += RB)
        for (int k = 0; k < 4; k++) t[k] += part[4 * b + k];
    for (int k = 0; k < 4; k++) t[k] = block_sum(t[k], sh);
    if (threadIdx.x == 0) {
        const double N = (double)n;
        const double num = N * t[3] - t[0] * t[1];
        const double den = N * t[2] - t[1] * t[1];
        const double a = num / den;
        const double b = t[0] / N - a * (t[1] / N);
        for (int k = 0; k < 4; k++) red[k] = t[k];
        red[4] = a;
        red[5] = b;
        if (closed) red[6] = red[10] - a * t[3] - b * t[0];
    }
}

// partial sums: r^2, r*G1, r*G2, r*G3 (G may be null)
__global__ void __launch_bounds__(RB) k_rse_resid(const float *__restrict__ Y,
                                                  const float *__restrict__ X,
                                                  const double *__restrict__ red,
                                                  const float *__restrict__ G1,
                                                  const float *__restrict__ G2,
                                                  const float *__restrict__ G3, long n,
                                                  double *__restrict__ part)
{
    __shared__ double sh[RB];
    const double a = red[4], b = red[5];
    double e = 0, g1 = 0, g2 = 0, g3 = 0;
    for (long i = (long)blockIdx.x * RB + threadIdx.x; i < n; i += (long)gridDim.x * RB) {
        const double r = (double)Y[i] - a * (double)X[i] - b;
        e += r * r;
        if (G1) {
            g1 += r * (double)G1[i];
            g2 += r * (double)G2[i];
            g3 += r * (double)G3[i];
        }
    }
    e = block_sum(e, sh);
    g1 = block_sum(g1, sh);
    g2 = block_sum(g2, sh);
    g3 = block_sum(g3, sh);
    if (threadIdx.x == 0) {
        double *p = part + 4 * blockIdx.x;
        p[0] = e;
        p[1] = g1;
        p[2] = g2;
        p[3] = g3;
    }
}

// red[6..9] = sum r^2, sum r G1, sum r G2, sum r G3
__global__ void __launch_bounds__(RB) k_rse_finish(const double *__restrict__ part, int nparts,
                                                   double *__restrict__ red)
{
    __shared__ double sh[RB];
    double t[4] = {0, 0, 0, 0};
    for (int b = threadIdx.x; b < nparts; b += RB)
        for (int k = 0; k < 4; k++) t[k] += part[4 * b + k];
    for (int k = 0; k < 4; k++) t[k] = block_sum(t[k], sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) red[6 + k] = t[k];
}

static unsigned grid_n(long n)
{
    long g = (n + RB - 1) / RB;
    return (unsigned)(g > 2048 ? 2048 : (g < 1 ? 1 : g));
}

hipError_t launch_rse_yy(const RsePlan &p, hipStream_t s)
{
    const unsigned gn = grid_n(p.n);
    hipLaunchKernelGGL(k_rse_sums, dim3(gn), dim3(RB), 0, s, p.Y, p.Y, p.n, p.part);
    hipLaunchKernelGGL(k_rse_yy, dim3(1), dim3(RB), 0, s, p.part, (int)gn, p.red);
    return hipGetLastError();
}

hipError_t launch_rse_prep(const RsePlan &p, const float *out4, hipStream_t s)
{
    if (p.mp) hipLaunchKernelGGL(k_rse_mpstats, dim3(p.M * p.M), dim3(RB), 0, s, out4, p.w, p.h,
                                 p.M, p.stats);
    hipLaunchKernelGGL(k_rse_prep, dim3(grid_n(p.n)), dim3(RB), 0, s, out4, p.stats, p.w, p.h,
                       p.M, p.crop, p.mp, p.Y, p.SR, p.Dt, p.DS);
    const hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_rse_yy(p, s);
}

hipError_t launch_rse_eval(const RsePlan &p, float sigma, bool grad, hipStream_t s)
{
    const int K = p.K;
    const float *g = p.taps, *dg = p.taps + 2 * K + 1;
    const unsigned gn = grid_n(p.n);
    if (!grad) {
        // value only: x = Gx Gy SR with the taps formed per block, the first-pass sums, and
        // alpha, beta and the squared residuals in closed form
        const size_t lt = sizeof(float) * (size_t)(2 * K + 1);
        hipLaunchKernelGGL(k_blur_cols_sig, dim3(gn), dim3(RB), lt, s, p.SR, p.T1, p.Wc, p.Hc,
                           sigma, K);
        hipLaunchKernelGGL(k_blur_rows_sig, dim3(gn), dim3(RB), lt, s, p.T1, p.X, p.Wc, p.Hc,
                           sigma, K);
        hipLaunchKernelGGL(k_rse_sums, dim3(gn), dim3(RB), 0, s, p.Y, p.X, p.n, p.part);
        hipLaunchKernelGGL(k_rse_alpha_beta, dim3(1), dim3(RB), 0, s, p.part, (int)gn, p.n,
                           p.red, 1);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_gauss_taps, dim3(1), dim3(RB), 0, s, sigma, K, p.taps);
    const size_t lt1 = sizeof(float) * (size_t)(2 * K + 1), lt2 = 2 * lt1;
    // x = Gx Gy SR (T1 = Gy SR)
    hipLaunchKernelGGL(k_blur_cols, dim3(gn), dim3(RB), lt1, s, p.SR, p.T1, p.Wc, p.Hc, g, K);
    hipLaunchKernelGGL(k_blur_rows, dim3(gn), dim3(RB), lt2, s, p.T1, g, (const float *)nullptr,
                       (const float *)nullptr, p.X, p.Wc, p.Hc, K);
    if (grad) {
        // dx/dsigma = dGx Gy SR + Gx dGy SR
        hipLaunchKernelGGL(k_blur_cols, dim3(gn), dim3(RB), lt1, s, p.SR, p.T2, p.Wc, p.Hc, dg, K);
        hipLaunchKernelGGL(k_blur_rows, dim3(gn), dim3(RB), lt2, s, p.T1, dg, p.T2, g, p.G3, p.Wc,
                           p.Hc, K);
        // blurred forward-mode derivatives of SR
        hipLaunchKernelGGL(k_blur_cols, dim3(gn), dim3(RB), lt1, s, p.Dt, p.T2, p.Wc, p.Hc, g, K);
        hipLaunchKernelGGL(k_blur_rows, dim3(gn), dim3(RB), lt2, s, p.T2, g, (const float *)nullptr,
                           (const float *)nullptr, p.G1, p.Wc, p.Hc, K);
        hipLaunchKernelGGL(k_blur_cols, dim3(gn), dim3(RB), lt1, s, p.DS, p.T2, p.Wc, p.Hc, g, K);
        hipLaunchKernelGGL(k_blur_rows, dim3(gn), dim3(RB), lt2, s, p.T2, g, (const float *)nullptr,
                           (const float *)nullptr, p.G2, p.Wc, p.Hc, K);
    }
    hipLaunchKernelGGL(k_rse_sums, dim3(gn), dim3(RB), 0, s, p.Y, p.X, p.n, p.part);
    hipLaunchKernelGGL(k_rse_alpha_beta, dim3(1), dim3(RB), 0, s, p.part, (int)gn, p.n, p.red,
                       0);
    hipLaunchKernelGGL(k_rse_resid, dim3(gn), dim3(RB), 0, s, p.Y, p.X, p.red,
                       grad ? p.G1 : (const float *)nullptr, grad ? p.G2 : (const float *)nullptr,
                       grad ? p.G3 : (const float *)nullptr, p.n, p.part);
    hipLaunchKernelGGL(k_rse_finish, dim3(1), dim3(RB), 0, s, p.part, (int)gn, p.red);
    return hipGetLastError();
}

int rse_max_parts(long n) { return (int)grid_n(n); }

}  // namespace esrrf
