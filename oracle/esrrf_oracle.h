/*
 * esrrf_oracle.h -- CPU restatement of the eSRRF path (test infrastructure only;
 * never linked into the product).  See esrrf_oracle.c for citations.
 */
#ifndef ESRRF_ORACLE_H
#define ESRRF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int width, height, magnification;
    float sensitivity;
    float radius, tSS, tSO;
    int nFrameForSRRF;
    int intWeighting;
} orc_bic_cfg;

typedef struct {
    int widthM, heightM; /* magnified frame size */
    float sensitivity;
    float radius, tSS, tSO;
    int nFrameForSRRF;
    int intWeighting;
} orc_fht_cfg;

typedef struct {
    int width, height, nPlanes, magnification;
    float sensitivity;
    float radius, tSS, tSO;
    int nFrameForSRRF;
    int intWeighting;
    int axialMagnification;   /* output planes per input plane; 0 = magnification */
} orc_3d_cfg;

void orc_java_float_to_string(float x, char *out, int len);
float orc_java_tso(float two_sigma);
void orc_bic_constants(float fwhm, int gradientMag, float *sigma, float *radius, float *tSS,
                       float *tSO);
void orc_fht_constants(float fwhm, int magnification, float *sigma, float *radius, float *tSS,
                       float *tSO);
float orc_cubic(float x);
float orc_interp(const float *a, int w, int h, float x, float y);

void orc_gradient_2point(const float *px, int w, int h, int nF, float *gx, float *gy);
void orc_gradient_interp(const float *g, int w, int h, int nF, float *gint);
void orc_bic_rgc_frame(const orc_bic_cfg *c, const float *px, const float *gxi,
                       const float *gyi, const float *shiftXY, int count0, int count1,
                       float *out, float *prev, int nthreads);
void orc_var(float *out, long n);
void orc_mp_map(const float *out, int w, int h, int M, float *mp);
void orc_mp_correct(float *out, int w, int h, int M, const float *mp);

void orc_gradient_robx(const float *px, int w, int h, int nF, float *gx, float *gy);
void orc_fht_rgc_frame(const orc_fht_cfg *c, const float *px, const float *gx,
                       const float *gy, const float *shiftXY, int count0, int count1,
                       float *out, float *prev, int nthreads);

void orc3d_align(const float *px, int w, int h, int nP, int nF, const float *st,
                 float *aligned);
void orc3d_gradient(const float *px, int w, int h, int nP, int nF, float *gx, float *gy,
                    float *gz);
void orc3d_interp(const float *g, int w, int h, int nP, int nF, float *gi);
void orc3d_rgc_frame(const orc_3d_cfg *c, const float *px, const float *gxi,
                     const float *gyi, const float *gzi, const float *shiftXY, int count0,
                     int count1, float *out, float *prev, int nthreads);
void orc3d_mp_map(const float *out, int w, int h, int nP, int M, int Mz, float *mp);
void orc3d_mp_correct(float *out, int w, int h, int nP, int M, int Mz, const float *mp);
void orc3d_reshape(const float *frame, int W, int H, int nP, const double *roi,
                   const double *intCoeff, float bg, float *planes);

#ifdef __cplusplus
}
#endif
#endif
