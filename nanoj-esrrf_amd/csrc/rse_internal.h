// rse_internal.h -- buffers and launches of the RSE value/gradient pipeline.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/esrrf.h"

namespace esrrf {

struct RsePlan {
    int w, h, M;           // raw frame, magnification
    int crop;              // SR pixels cropped on every side (cropSize * M)
    int Wc, Hc;            // cropped size
    long n;                // Wc * Hc
    int mp;                // MP-correct the AVG image
    int K;                 // blur half-width (fixed per problem)
    double *stats;         // [3][M*M] means of AVG, dAVG/dtSS, dAVG/dS per sub-position
    float *Y, *SR, *Dt, *DS;      // cropped ref, SR and its derivatives
    float *T1, *T2, *X, *G1, *G2, *G3;   // blur scratch and blurred images
    float *taps;           // [2][2K+1]
    double *part;          // [rse_max_parts(n)][4]
    double *red;           // [12]: sums, alpha, beta, E, E-gradient dot products, sum y^2
};

hipError_t launch_rse_prep(const RsePlan &p, const float *out4, hipStream_t s);
// sum y^2 of the cropped reference into red[10] (after the reference changes)
hipError_t launch_rse_yy(const RsePlan &p, hipStream_t s);
// value (grad = false) or value and the dot products of the gradient (grad = true)
hipError_t launch_rse_eval(const RsePlan &p, float sigma, bool grad, hipStream_t s);
int rse_max_parts(long n);

// Brent's derivative-free minimisation (R. P. Brent, Algorithms for Minimization without
// Derivatives, 1973, ch. 5), as commons-math3's BrentOptimizer applies it for the RSF
// sigma (ErrorMapLiveSRRF.java:119-127: relative 1e-10, absolute 1e-14, 1000 evaluations).
template <class F>
inline int brent(F f, double lo, double hi, double *xbest, double *fbest, int *nev)
{
    const double c = 0.5 * (3.0 - std::sqrt(5.0));
    const double rel = 1e-10, abs_ = 1e-14;
    double a = lo, b = hi;
    double x = a + c * (b - a), w = x, v = x;
    double fx, fw, fv;
    int rc = f(x, &fx);
    if (rc) return rc;
    fw = fv = fx;
    int n = 1;
    double d = 0.0, e = 0.0;
    while (n < 1000) {
        const double m = 0.5 * (a + b);
        const double tol = rel * std::fabs(x) + abs_, t2 = 2.0 * tol;
        if (std::fabs(x - m) <= t2 - 0.5 * (b - a)) break;
        double p = 0.0, q = 0.0, r = 0.0;
        if (std::fabs(e) > tol) {
            r = (x - w) * (fx - fv);
            q = (x - v) * (fx - fw);
            p = (x - v) * q - (x - w) * r;
            q = 2.0 * (q - r);
            if (q > 0) p = -p;
            else q = -q;
            r = e;
            e = d;
        }
        if (std::fabs(p) < std::fabs(0.5 * q * r) && p > q * (a - x) && p < q * (b - x)) {
            d = p / q;
            const double u = x + d;
            if (u - a < t2 || b - u < t2) d = x <= m ? tol : -tol;
        } else {
            e = x < m ? b - x : a - x;
            d = c * e;
        }
        const double u = x + (std::fabs(d) >= tol ? d : (d > 0 ? tol : -tol));
        double fu;
        rc = f(u, &fu);
        if (rc) return rc;
        n++;
        if (fu <= fx) {
            if (u < x) b = x;
            else a = x;
            v = w;
            fv = fw;
            w = x;
            fw = fx;
            x = u;
            fx = fu;
        } else {
            if (u < x) a = u;
            else b = u;
            if (fu <= fw || w == x) {
                v = w;
                fv = fw;
                w = u;
                fw = fu;
            } else if (fu <= fv || v == x || v == w) {
                v = u;
                fv = fu;
            }
        }
    }
    *xbest = x;
    *fbest = fx;
    if (nev) *nev = n;
    return ESRRF_OK;
}



}  // namespace esrrf

namespace esrrf {

// SQUIRREL error map of a (ref, SR) pair (k_squirrel.hip); p.Y / p.SR receive the crops.
hipError_t launch_squirrel_crop(const float *ref, const float *sr, int W, int crop,
                                const RsePlan &p, hipStream_t s);
// after launch_rse_eval at the chosen sigma: error map and RSC (each may be null), then
// out[0..3] = RSE, RSP, alpha, beta; part holds squirrel_max_parts(p.n) * 6 doubles
hipError_t launch_squirrel_maps(const RsePlan &p, float *emap, float *rsc, double *part,
                                double *out, hipStream_t s);
int squirrel_max_parts(long n);

}  // namespace esrrf
