/*
 * esrrf.h -- C ABI of the MI355X eSRRF reconstruction engine (libesrrf.so).
 *
 * This is the boundary the Fiji plugin's host class LiveSRRF_CL
 * (src/nanoj/liveSRRF/LiveSRRF_CL.java) binds through Java FFM or JNI in place of its
 * JOCL/OpenCL calls.  Every entry point below names the LiveSRRF_CL method it replaces.
 * Plain pointers and sizes only; the caller owns host memory, the engine owns device
 * memory.  All calls return 0 (ESRRF_OK) or a positive ESRRF_ERR_* code; the message of
 * the last failure of an engine is esrrf_last_error().  One engine may be driven from
 * several threads: every call takes the engine's mutex (LiveSRRF_CL.calculateSRRF is
 * `synchronized`, LiveSRRF_CL.java:636).
 */
#ifndef ESRRF_H
#define ESRRF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ESRRF_OK 0
#define ESRRF_ERR_INVALID 1     /* bad argument or size */
#define ESRRF_ERR_HIP 2         /* HIP runtime failure (message has the HIP error) */
#define ESRRF_ERR_STATE 3       /* call not valid in the engine's current state */
#define ESRRF_ERR_NOMEM 4       /* device allocation failed */
#define ESRRF_ERR_UNSUPPORTED 5 /* configuration outside what the engine implements */

#define ESRRF_U16 0 /* frames as uint16 (ImageJ 16-bit stacks) */
#define ESRRF_F32 1 /* frames as float32 */

typedef struct esrrf_engine esrrf_engine;

/* Arguments of LiveSRRF_CL.initialise (LiveSRRF_CL.java:159).  The 3-D calibration table
 * that initialise reads from calibTablePath3DSRRF (:207-227) is passed as arrays: set
 * n_planes > 0 to select 3-D (MFM) mode, which like the reference forces BIC
 * interpolation (:166-168). */
typedef struct esrrf_params {
    int width;                /* raw frame width (3-D: full 3x3 montage width)   */
    int height;               /* raw frame height (3-D: full montage height)     */
    int magnification;
    float fwhm;               /* the GUI's "Radius" (gui/LiveSRRF_optimised_.java:606) */
    float sensitivity;
    int n_frames_on_gpu;      /* frames per calculate call (one GPU load)        */
    int n_frame_for_srrf;     /* frames per SR frame                              */
    int block_length;         /* accepted for API parity; the engine tiles itself */
    int device;               /* HIP device index; -1 = max CUs x clock (:121-137) */
    int int_weighting;
    int do_mp_map_correction;
    int do_fht_interpolation;
    int n_planes;             /* 0 = 2-D; >0 = 3-D with this many planes          */
    const double *shift_x_3d; /* [n_planes] "X-shift (pixels)"                   */
    const double *shift_y_3d; /* [n_planes] "Y-shift (pixels)"                   */
    const double *theta_3d;   /* [n_planes] "Theta (degrees)"                    */
    const double *roi_3d;     /* [n_planes] "ROI #"                              */
    const double *int_coeffs_3d; /* [n_planes] "Intensity scaling"               */
    float background_3d;      /* "Background level (ADC)"[0]                     */
    float axial_offset;       /* nm (:236)                                       */
    float pixel_size;         /* um (:236)                                       */
    int axial_magnification;  /* 3-D: output planes per input plane; 0 = magnification, the
                               * reference's single magnification (live3DSRRF.cl:354-356).
                               * No counterpart: the axial centre is (zM + 0.5) / this value */
} esrrf_params;

/* LiveSRRF_CL.checkDevices (LiveSRRF_CL.java:111-155): number of HIP devices. */
int esrrf_device_count(int *n);
/* CLDevice.getName() of the device list built by checkDevices (:126-153). */
int esrrf_device_name(int device, char *buf, int len);
/* Device with the largest compute units x max clock product (:121-137). */
int esrrf_fastest_device(int *device);

/* LiveSRRF_CL.initialise (:159-264) + prepareBuffersBICinterpolation (:266-459) /
 * prepareBuffersFHTinterpolation (:462-579): allocate, derive constants, build tables. */
int esrrf_create(const esrrf_params *p, esrrf_engine **out);
/* LiveSRRF_CL.release (:939-951). */
int esrrf_destroy(esrrf_engine *e);

/* LiveSRRF_CL.loadDriftXYGPUbuffer (:583-595): drift in raw pixels, stored x M. */
int esrrf_load_drift(esrrf_engine *e, const float *drift_x, const float *drift_y, int n);
/* LiveSRRF_CL.resetFramePosition (:955-960): start a new SR frame. */
int esrrf_reset_frame_position(esrrf_engine *e);

/* LiveSRRF_CL.calculateSRRF (:636-746): push n_frames host frames (one GPU load) of
 * width x height (dtype ESRRF_U16 / ESRRF_F32), run gradients + RGC + accumulation.
 * `cancel` (may be NULL) is polled between launch chunks, like IJ.escapePressed()
 * (:731-734); *cancelled receives the method's boolean result. */
int esrrf_calculate(esrrf_engine *e, const void *frames, int dtype, int n_frames,
                    const volatile int *cancel, int *cancelled);
/* Same with frames already resident in device memory (no counterpart: the reference
 * always uploads from the JVM heap). */
int esrrf_calculate_device(esrrf_engine *e, const void *d_frames, int dtype, int n_frames,
                           const volatile int *cancel, int *cancelled);

/* LiveSRRF_CL.prepareDataFHT_SRRF (:597-633), FHT mode only: upload n_frames host frames,
 * Fourier-upsample them by the magnification (a plain copy at M = 1) and form their
 * gradients, without the RGC; resets the load counter like kernelCalculateGradient.
 * esrrf_read_gradients then returns frame 0's Gx, Gy.  ESRRF_ERR_STATE in other modes. */
int esrrf_prepare_fht(esrrf_engine *e, const void *frames, int dtype, int n_frames);
/* LiveSRRF_CL.reshapeImageStack3D (:963-994), 3-D mode only: n_frames host montage frames
 * (width x height) -> out [n_frames][n_planes][height_s][width_s] float, each plane the
 * calibration's montage cell, (p - background) / intensity scaling, computed on the
 * device.  Leaves the SR state unchanged.  ESRRF_ERR_STATE in other modes. */
int esrrf_reshape_3d(esrrf_engine *e, const void *frames, int dtype, int n_frames, float *out);

/* LiveSRRF_CL.readSRRFbuffer (:749-794): VAR/TAC2 finalise, MP correction (BIC and 3-D
 * only, :762), NaN -> 0, copy out [4][n_planes_m][hM][wM] (AVG, VAR, TAC2, INT). */
int esrrf_read_result(esrrf_engine *e, float *out);
int esrrf_read_result_device(esrrf_engine *e, float *d_out);

/* LiveSRRF_CL.readGradientBuffers (:797-878): frame 0 of the current load; BIC/3-D: per
 * plane GxInt, GyInt[, GzInt] of (2w x 2h); FHT: Gx, Gy of (wM x hM). */
int esrrf_read_gradients(esrrf_engine *e, float *out);
/* LiveSRRF_CL.readAlignedPixels (:881-905): 3-D aligned planes of frame 0. */
int esrrf_read_aligned(esrrf_engine *e, float *out);
/* LiveSRRF_CL.readMPmaps (:909-935): [3 (x M in 3-D)][M][M]. */
int esrrf_read_mp_maps(esrrf_engine *e, float *out);
/* The nCurrentFrame[2] buffer (liveSRRF.cl:245-248): [0] frame in SR frame, [1] frame in
 * load. */
int esrrf_read_frame_counters(esrrf_engine *e, int *out2);
/* Public fields widthS, heightS, nPlanes, nPlanesM (:41-43) and the magnified size. */
int esrrf_get_dims(esrrf_engine *e, int *width_m, int *height_m, int *n_planes_m,
                   int *width_s, int *height_s, int *n_planes);
/* LiveSRRF_CL.finishQueue (:1009-1015). */
int esrrf_finish(esrrf_engine *e);

/* Run on the caller's HIP stream (NULL restores the engine's own stream). */
int esrrf_set_stream(esrrf_engine *e, void *hip_stream);
/* esrrf_calculate (BIC mode, no cancel flag) splits a load of host frames into up to
 * `chunks` (1..16, at most one per 64 frames) parts, each three times the one before (the
 * first at least 16 frames), and copies part c+1 while part c is computed, so that only the
 * first, small copy precedes the kernels.  No counterpart: the reference queues one write
 * of the whole load ahead of the kernels on its in-order queue (LiveSRRF_CL.java:655-661).
 * 1 = one copy, then the kernels.  Default 4. */
int esrrf_set_upload_chunks(esrrf_engine *e, int chunks);
/* Page-locked host memory from the runtime the engine copies with, for frame and result
 * buffers that upload and read back at the link's full rate (no counterpart: the reference
 * copies from the JVM heap).  Pageable memory works everywhere, more slowly. */
int esrrf_host_alloc(size_t bytes, void **ptr);
int esrrf_host_free(void *ptr);
/* Kernel timing with HIP events on the launch stream (NanoJProfiler equivalent,
 * LiveSRRF_CL.java:67).  which: 0 = RGC kernel, 1 = gradient prep, 2 = upsampling. */
int esrrf_profile_enable(esrrf_engine *e, int enable);
int esrrf_profile_read(esrrf_engine *e, int which, double *total_ms, long *launches);

/* Constants of prepareBuffersBICinterpolation (:304-305,317-320).  tSO is the float the
 * kernel reads from the text `"" + 2 * sigma + 1` (:318), which Java evaluates as string
 * concatenation: Float.toString(2*sigma) followed by "1" (1.27442651 at fwhm 1.5). */
int esrrf_bic_constants(float fwhm, float *sigma, float *radius, float *tSS, float *tSO);
/* The $TWOSIGpONE$ text of :318 itself, NUL-terminated, into buf[len]. */
int esrrf_bic_tso_text(float fwhm, char *buf, int len);
/* Constants of prepareBuffersFHTinterpolation (:491-500). */
int esrrf_fht_constants(float fwhm, int magnification, float *sigma, float *radius,
                        float *tSS, float *tSO);
/* Gated RGC sample count of one frame (geometry only) -- the algorithmic work unit. */
int esrrf_active_samples(esrrf_engine *e, long long *per_frame);

/* Rolling-analysis bookkeeping of gui/LiveSRRF_optimised_.java:621-647, 842-877.
 * out6 = {nFrameForSRRF, frameGap, nSRRFframe, nFrameOnGPU, nGPUloadPerSRRFframe, 0}. */
int esrrf_plan_rolling(int n_slices, int n_frame_for_srrf_user, int frame_gap_user,
                       int rolling, int max_frames_on_gpu, int *out6);
/* Load l of SR frame r (1-based) of that plan: first stack index (1-based) and size
 * (gui/LiveSRRF_optimised_.java:281,311-321). */
int esrrf_plan_load(const int *plan6, int r, int l, int *first_index, int *n_frames);

/* Autotuner of the tiled 2-D kernel (no counterpart in the reference, whose launch shape is
 * the GUI's blockLength, LiveSRRF_CL.java:699-735): times every supported launch shape
 * (waves per block x tap tables in LDS or through scalar loads x one or two frames per
 * pass) over the frames of the current load, `reps`
 * launches each, into scratch outputs, and keeps the fastest for this engine.  Reports it. */
int esrrf_autotune(esrrf_engine *e, int reps, int *nw, int *ppw, int *taps_in_lds, float *ms);
/* Launch shape the engine's tiled 2-D kernel uses now (no counterpart in the reference):
 * waves per block, sub-pixel phases per wave, tap tables in LDS, and frames per pass
 * (2 or 4: two or four frames share each tap read).  ESRRF_ERR_STATE when the engine runs
 * no tiled kernel (FHT, 3-D, or a geometry outside the tiled kernel's range). */
int esrrf_launch_shape(esrrf_engine *e, int *nw, int *ppw, int *taps_in_lds,
                       int *frames_per_pass);
/* Pin the tiled 2-D kernel's launch shape (no counterpart in the reference), e.g. one the
 * autotuner chose earlier: nw waves per block (1..16, at most 8 sub-pixel phases per wave),
 * tap tables in LDS (1) or through scalar loads (0), one, two or four frames per pass.
 * ESRRF_ERR_UNSUPPORTED when the kernels cannot run that shape for this geometry (tables or
 * tiles beyond their LDS budget, two or four frames with more than 4 phases per wave, four
 * frames with the tables outside LDS or a tile wider than that kernel's);
 * ESRRF_ERR_STATE when the engine runs no tiled kernel. */
int esrrf_set_launch_shape(esrrf_engine *e, int nw, int taps_in_lds, int frames_per_pass);
/* Name of the RGC kernel the engine's last calculate launched, e.g. "k_bic_rgc4",
 * "k_fht_rgc2", "k3d_rgc3" ("" before the first launch).  No counterpart in the reference
 * (its single OpenCL kernel per mode needs none); for profiling and tests. */
int esrrf_rgc_kernel(esrrf_engine *e, char *buf, int len);
/* ---- Parameter sweep (gui/ParametersSweep_.java:328-519) --------------------------------
 * Every (magnification, frame count, sensitivity, fwhm) configuration of a sweep over the
 * first nf frames of one stack, 2-D BIC, no drift; for each the four channels that
 * readSRRFbuffer returns after that configuration's initialise/calculateSRRF/readSRRFbuffer
 * (LiveSRRF_CL.java:159,636,749), computed from one upload, one gradient pass, one RGC
 * launch per magnification and group of up to five fwhm values (the fwhm only changes the
 * weights and the gate of a sample, so one angular term serves the group) and one
 * accumulation launch per fwhm for all frame counts and sensitivities.  The reference loops nf -> S -> fwhm and rebuilds its program per
 * configuration (:328-380, :452). */
typedef struct esrrf_sweep esrrf_sweep;
typedef struct esrrf_sweep_params {
    int width, height;                  /* raw frame size                                */
    const int *magnifications;          /* 1..8 each                                      */
    int n_magnifications;
    const float *fwhm;                  /* the GUI's "Radius" values (:720-723)           */
    int n_fwhm;
    const float *sensitivity;           /* (:725-728), at most 32 values                  */
    int n_sensitivity;
    const int *n_frames;                /* frames per SR image, first n frames (:730-747), at most 8 */
    int n_n_frames;
    int int_weighting;                  /* the reference sweep passes 1 (:452)            */
    int do_mp_map_correction;           /* the reference sweep passes 1 (:452)            */
    int autotune;                       /* 1: time the RGC launch shapes per magnification */
    int use_graph;                      /* 1: capture the device sequence in a hipGraph   */
    int device;                         /* -1: fastest                                    */
    int do_fht_interpolation;           /* 1: FHT mode, the reference sweep's default
                                         * (:117,452): frames upsampled once per
                                         * magnification, no MP correction (:762)         */
} esrrf_sweep_params;
int esrrf_sweep_create(const esrrf_sweep_params *p, esrrf_sweep **out);
int esrrf_sweep_destroy(esrrf_sweep *s);
int esrrf_sweep_set_stream(esrrf_sweep *s, void *hip_stream);
/* Run every configuration on the first max(n_frames) frames of a host stack (blocking). */
int esrrf_sweep_run(esrrf_sweep *s, const void *frames, int dtype, int n_frames);
/* Same from device memory, asynchronous on the sweep's stream. */
int esrrf_sweep_run_device(esrrf_sweep *s, const void *d_frames, int dtype, int n_frames);
/* One configuration's [4][hM][wM] result (AVG, VAR, TAC2, INT; NaN -> 0), indices into the
 * parameter lists. */
int esrrf_sweep_read(esrrf_sweep *s, int i_mag, int i_nf, int i_sens, int i_fwhm, float *out);
int esrrf_sweep_read_device(esrrf_sweep *s, int i_mag, int i_nf, int i_sens, int i_fwhm,
                            float *d_out);
/* Launch shape of the emit launch that serves a (magnification, fwhm) (autotuned or
 * default): waves per block, phases per wave, tap tables in LDS, frames per pass, autotuner
 * time (0: not tuned). */
int esrrf_sweep_shape(esrrf_sweep *s, int i_mag, int i_fwhm, int *nw, int *ppw, int *taps_in_lds,
                      int *frames_per_pass, float *ms);
const char *esrrf_sweep_last_error(esrrf_sweep *s);

/* ---- SQUIRREL RSE and its gradient (ErrorMapLiveSRRF.java:81-246) ------------------------
 * RSE of the eSRRF AVG image of a loaded stack (2-D BIC, no drift, one SR image over all
 * loaded frames) against its INT image, as the parameter sweep scores a configuration
 * (gui/ParametersSweep_.java:473-476): crop crop_size*M SR pixels per side, blur with a
 * Gaussian RSF of sigma SR pixels, alpha/beta least squares, RSE = sqrt(mean(residual^2)).
 * The blur is a separable Gaussian truncated at ceil(5 sigma_max) with clamp-to-edge borders
 * (stand-in for ImageJ's blurGaussian).  The gradient with respect to (fwhm, sensitivity,
 * sigma) comes from an adjoint variant of the RGC kernel; no counterpart in the reference. */
typedef struct esrrf_rse esrrf_rse;
typedef struct esrrf_rse_params {
    int width, height;        /* raw frame size                                             */
    int magnification;        /* 1..8                                                       */
    int n_frames;             /* largest stack to load (frames per SR image)                */
    int crop_size;            /* raw pixels cropped per side, cropSize (default 3 in the GUI) */
    int int_weighting;        /* 1 as the reference sweep                                   */
    int do_mp_map_correction; /* 1 as the reference sweep                                   */
    float sigma_max;          /* RSF sigma bound in SR pixels; <= 0: 4/2.35482*M (:75)      */
    int device;               /* -1: fastest                                                */
} esrrf_rse_params;
int esrrf_rse_create(const esrrf_rse_params *p, esrrf_rse **out);
int esrrf_rse_destroy(esrrf_rse *r);
/* Upload a stack (2 <= n <= n_frames) and compute its unit tap gradients. */
int esrrf_rse_load(esrrf_rse *r, const void *frames, int dtype, int n_frames);
/* RSE at (fwhm, sensitivity, sigma); grad3 (may be NULL) receives dRSE/d(fwhm, S, sigma). */
int esrrf_rse_value_and_grad(esrrf_rse *r, float fwhm, float sensitivity, float sigma,
                             double *rse, double *grad3);
/* The sigma that minimises the RSE on [0, sigma_max] (Brent, ErrorMapLiveSRRF.java:119-127). */
int esrrf_rse_best_sigma(esrrf_rse *r, float fwhm, float sensitivity, double *sigma,
                         double *rse, int *n_eval);
/* Projected Adam descent on (fwhm, S, sigma) from x0[3] for `steps` gradient steps;
 * lr/lo/hi (each [3], may be NULL for defaults) are step sizes and box bounds;
 * trajectory[(steps+1)*4] receives (fwhm, S, sigma, RSE) of every evaluated point. */
int esrrf_rse_descend(esrrf_rse *r, const float *x0, int steps, const float *lr,
                      const float *lo, const float *hi, double *trajectory);
/* Evaluations and RGC passes made so far (each new (fwhm, S) costs one RGC pass). */
int esrrf_rse_counters(esrrf_rse *r, long long *evaluations, long long *rgc_passes);
const char *esrrf_rse_last_error(esrrf_rse *r);

/* ---- SQUIRREL error map of any (reference, SR) image pair (ErrorMapLiveSRRF.java:50-193) --
 * optimise(impRef, impSR) then calculateErrorMap() on device images of (width*M) x (height*M)
 * pixels: crop crop_size*M SR pixels per side when crop_border (:57-64, :88-99), find the RSF
 * sigma by Brent on [0, 4/2.35482*M] (:119-127; or take fixed_sigma*M, :53,:116-117), fit
 * alpha, beta by least squares of ref on (blur(SR), blur(1)) (:145-152, :201-233), and form
 * RSC = blur(alpha SR + beta), RSE = sqrt(mean((RSC - ref)^2)) (globalRMSE), RSP = Pearson
 * correlation of RSC and ref (globalPPMCC) and the error map |ref - RSC| (:160-190).  NaN
 * pixels of the inputs read as 0 (readSRRFbuffer hands over NaN -> 0 images).  The blur is a
 * separable Gaussian truncated at ceil(5 sigma_max) with clamp-to-edge borders (stand-in for
 * ImageJ's blurGaussian), and sums run in double. */
typedef struct esrrf_squirrel esrrf_squirrel;
typedef struct esrrf_squirrel_params {
    int width, height;        /* raw frame size (imp of ErrorMapLiveSRRF's constructor)      */
    int magnification;        /* 1..16                                                      */
    int crop_border;          /* 1: crop crop_size raw pixels per side (cropBorder)          */
    int crop_size;            /* cropSize (3 in the sweep, gui/ParametersSweep_.java:114)    */
    int fix_sigma;            /* 1: RSF sigma = fixed_sigma * M (fixSigma, fixedSigma)       */
    float fixed_sigma;        /* raw pixels                                                 */
    int device;               /* -1: fastest                                                */
} esrrf_squirrel_params;
typedef struct esrrf_squirrel_result {
    double sigma;             /* RSF sigma in SR pixels (sigma_linear)                      */
    double alpha, beta;
    double rse;               /* globalRMSE                                                 */
    double rsp;               /* globalPPMCC                                                */
    int overblur;             /* 1: sigma within 1e-4 of its bound (OVERBLUR_MESSAGE, :135) */
    int evaluations;          /* objective evaluations of the sigma search                  */
} esrrf_squirrel_result;
int esrrf_squirrel_create(const esrrf_squirrel_params *p, esrrf_squirrel **out);
int esrrf_squirrel_destroy(esrrf_squirrel *q);
int esrrf_squirrel_set_stream(esrrf_squirrel *q, void *hip_stream);
/* Cropped size of the maps (Wc x Hc). */
int esrrf_squirrel_dims(esrrf_squirrel *q, int *width_c, int *height_c);
/* ref, sr: [height*M][width*M] device images; error_map, rsc: [Hc][Wc] device outputs or
 * NULL.  Blocking (the sigma search reads back a few doubles per evaluation). */
int esrrf_squirrel_run_device(esrrf_squirrel *q, const float *d_ref, const float *d_sr,
                              esrrf_squirrel_result *res, float *d_error_map, float *d_rsc);
/* Same from and to host memory. */
int esrrf_squirrel_run(esrrf_squirrel *q, const float *ref, const float *sr,
                       esrrf_squirrel_result *res, float *error_map, float *rsc);
const char *esrrf_squirrel_last_error(esrrf_squirrel *q);
/* The sweep's error mapping (gui/ParametersSweep_.java:465-517): for every configuration of
 * the last esrrf_sweep_run, SR = channel (0 AVG, 1 VAR, 2 TAC2) and ref = INT of that
 * configuration, cropped by crop_size raw pixels.  rse, rsp, sigma (each may be NULL) receive
 * [n_magnifications][n_n_frames][n_sensitivity][n_fwhm] values (the reference's RSE and RSP
 * sweep maps are the [n_sensitivity][n_fwhm] slices, :518-530). */
int esrrf_sweep_squirrel(esrrf_sweep *s, int channel, int crop_size, int fix_sigma,
                         float fixed_sigma, double *rse, double *rsp, double *sigma);
/* QnR map (gui/GetQnRmap_.java:179-234): each FRC resolution value r becomes
 * L(r) = 1/(exp(-k (r - x0)) + 1), the logistic through (res_min, 0.925) and (res_max, 0.075)
 * (:179-203); then QnR = 2 L RSP / (L + RSP) (:227-234).  With auto_min_max the bounds are the
 * smallest and largest finite FRC values (:86-90).  n values; host arrays.  The best value and
 * its index (first maximum) go to *best and *best_index (each may be NULL). */
int esrrf_qnr_map(const float *frc, const float *rsp, long n, int auto_min_max, float res_min,
                  float res_max, float *qnr, float *best, long *best_index);
/* One pass of ImageJ's "Remove NaNs..." on a [height][width] map, in place, as GetQnRmap_
 * runs it on the RSP and FRC maps (gui/GetQnRmap_.java:69-77, radius 2): every NaN pixel
 * becomes the median of the non-NaN pixels of the circular kernel dx^2 + dy^2 <= r^2 + 1
 * (edge pixels extend the map; values read before the pass); a pixel with no such
 * neighbour stays NaN.  ImageJ is not available: the kernel shape and the even-count median
 * (the mean of the two middle values) are the build's, parity unpinned.  Host code. */
int esrrf_remove_nans(float *map, int width, int height, double radius);

/* ---- Fourier ring correlation (the sweep's FRC resolution, gui/ParametersSweep_.java:403-447)
 * FRC.calculateFireNumber(ip1, ip2, FIXED_1_OVER_7) belongs to NanoJ-Core, which the reference
 * does not contain; this is the standard FRC: centred square crop, Tukey(0.25) taper, zero
 * padding to the power of two N >= the square's side (N <= 4096), ring r = nearest integer to
 * |k|, FRC(r) = sum Re(F1 conj F2) / sqrt(sum |F1|^2 sum |F2|^2) in double, resolution
 * N / r* in pixels at the first crossing r* of 1/7 after ring 0 (linear between rings), NaN
 * when the curve never crosses.  NaN pixels read as 0. */
typedef struct esrrf_frc esrrf_frc;
int esrrf_frc_create(int width, int height, int device, esrrf_frc **out);
int esrrf_frc_destroy(esrrf_frc *q);
int esrrf_frc_set_stream(esrrf_frc *q, void *hip_stream);
/* Rings of the curve (N/2) and the transform size N. */
int esrrf_frc_rings(esrrf_frc *q, int *n_rings, int *transform_size);
/* Two [height][width] device images; *resolution in pixels; curve (host, N/2 values) may be
 * NULL.  Blocking. */
int esrrf_frc_run_device(esrrf_frc *q, const float *d_a, const float *d_b, double *resolution,
                         float *curve);
int esrrf_frc_run(esrrf_frc *q, const float *a, const float *b, double *resolution,
                  float *curve);
const char *esrrf_frc_last_error(esrrf_frc *q);
/* No counterpart in the reference: esrrf_frc_run_device for callers that cannot allocate
 * device memory (code under stream capture that must not call hipMalloc, tight serving
 * loops, harnesses that forbid allocation).  No handle and no allocation: the spectra, ring
 * sums, window and twiddles live in one static device array of about 76 MiB per GPU, held
 * from the library's first use on that GPU to the end of the process.  Transform size N
 * (the power of two >= min(width, height)) at most ESRRF_FRC_NOALLOC_MAX_N, otherwise
 * ESRRF_ERR_UNSUPPORTED.  device < 0: the current device; hip_stream NULL: the null stream.
 * curve (host) receives N/2 values when given, *n_rings N/2.  Blocking.  Calls that overlap
 * on one GPU share the array and race: serialise them. */
#define ESRRF_FRC_NOALLOC_MAX_N 2048
int esrrf_frc_run_device_noalloc(int device, void *hip_stream, const float *d_a,
                                 const float *d_b, int width, int height, double *resolution,
                                 float *curve, int *n_rings);
/* FRC resolution (SR pixels) of every configuration of two sweeps with identical parameters
 * run on the odd and the even frames (:403-447 with :808-816's frame mapping), for channel 0
 * AVG, 1 VAR or 2 TAC2: [n_magnifications][n_n_frames][n_sensitivity][n_fwhm]. */
int esrrf_sweep_frc(esrrf_sweep *odd, esrrf_sweep *even, int channel, double *resolution);

/* ---- Drift of a window's frames (XYShiftCalculator.calculateShiftArray,
 * XYShiftCalculator.java:25-47, whose shiftX / shiftY feed loadDriftXYGPUbuffer,
 * gui/LiveSRRF_optimised_.java:284-305) ------------------------------------------------------
 * The reference reads the shift from NanoJ-Core's cross-correlation map and peak fit, which
 * the reference does not contain; this is the standard method: both frames minus their means,
 * zero-padded to the power of two N >= max(width, height) (N <= 4096), circular
 * cross-correlation CC(t) = sum_x r(x) f(x + t) by FFT, the first maximum (row-major), lags
 * of N/2 and above wrapped to negative, and per axis the parabola through the peak and its
 * two neighbours.  shift = the displacement of a frame's content relative to the reference
 * (f(x) = r(x - d) gives d), in raw pixels, the sign loadDriftXYGPUbuffer takes.  NaN pixels
 * read as 0. */
typedef struct esrrf_drift esrrf_drift;
int esrrf_drift_create(int width, int height, int device, esrrf_drift **out);
int esrrf_drift_destroy(esrrf_drift *q);
int esrrf_drift_set_stream(esrrf_drift *q, void *hip_stream);
/* Shifts of n_frames [height][width] device frames against frame ref_index (the reference
 * uses the window's first frame, :27); shift_x, shift_y (host, n_frames each; either may be
 * NULL).  Blocking. */
int esrrf_drift_run_device(esrrf_drift *q, const float *d_frames, int n_frames, int ref_index,
                           float *shift_x, float *shift_y);
int esrrf_drift_run(esrrf_drift *q, const float *frames, int n_frames, int ref_index,
                    float *shift_x, float *shift_y);
const char *esrrf_drift_last_error(esrrf_drift *q);

/* ---- Rolling analysis of a long stack (gui/LiveSRRF_optimised_.java:276-384) ------------
 * SR image r (0-based) covers frames [r*gap, r*gap + nF) of an n_slices stack, with nF, gap
 * and the number of SR images from the plugin's plan (:621-627, esrrf_plan_rolling with
 * rolling on); 2-D BIC, no drift.  Each result equals what initialise / calculateSRRF over
 * that window / readSRRFbuffer returns (LiveSRRF_CL.java:159,636,749).  Frames stream
 * through in chunks: one gradient and one RGC pass per frame serve every window containing
 * it, so overlapping windows (gap < nF) do not repeat the RGC.  When more than about 60
 * windows overlap one frame (gap < nF/61) esrrf_rolling_begin returns ESRRF_ERR_UNSUPPORTED
 * and callers loop the engine per window, as the plugin does. */
typedef struct esrrf_rolling esrrf_rolling;
typedef struct esrrf_rolling_params {
    int width, height;        /* raw frame size                                             */
    int magnification;        /* 1..8                                                       */
    float fwhm, sensitivity;
    int n_frame_for_srrf;     /* the user's value: 0 = every frame (:621-623)               */
    int frame_gap;            /* the user's value: 0 = n_frame_for_srrf (:624-627)          */
    int int_weighting;
    int do_mp_map_correction;
    int chunk_frames;         /* frames per RGC pass, <= 0: 256                             */
    int device;               /* -1: fastest                                                */
} esrrf_rolling_params;
int esrrf_rolling_create(const esrrf_rolling_params *p, esrrf_rolling **out);
int esrrf_rolling_destroy(esrrf_rolling *r);
int esrrf_rolling_set_stream(esrrf_rolling *r, void *hip_stream);
/* Start a run over n_slices frames (u16 or f32) in host memory, which must stay valid until
 * the run's last esrrf_rolling_next; *n_sr receives the number of SR images (nSRRFframe). */
int esrrf_rolling_begin(esrrf_rolling *r, const void *frames, int dtype, int n_slices,
                        int *n_sr);
/* Same with the stack in device memory. */
int esrrf_rolling_begin_device(esrrf_rolling *r, const void *d_frames, int dtype, int n_slices,
                               int *n_sr);
/* Compute until the next SR image (in order) is complete and copy its [4][hM][wM] (AVG, VAR,
 * TAC2, INT; NaN -> 0) to host memory; *index receives its number.  Blocking. */
int esrrf_rolling_next(esrrf_rolling *r, float *out, int *index);
/* Same into device memory, asynchronous on the rolling stream. */
int esrrf_rolling_next_device(esrrf_rolling *r, float *d_out, int *index);
const char *esrrf_rolling_last_error(esrrf_rolling *r);

const char *esrrf_last_error(esrrf_engine *e);
const char *esrrf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* ESRRF_H */
