// host_plan.cpp -- the engine's pure host arithmetic: the kernel constants LiveSRRF_CL bakes
// into its kernel text and the rolling/load plan of the plugin's frame loop.  No HIP calls,
// so the CPU sanitizer build links this file as it is.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>

#include "../../include/esrrf.h"
#include "java_text.h"

using namespace esrrf;

extern "C" {

int esrrf_bic_constants(float fwhm, float *sigma, float *radius, float *tSS, float *tSO)
{
    // LiveSRRF_CL.java:304-305,317-318 (float arithmetic, left to right).  `"" + 2 * sigma
    // + 1` (:318) is string concatenation: the kernel reads tSO from the text
    // Float.toString(2*sigma) + "1" as a double literal and casts it to float.
    volatile float s = fwhm / 2.354f;
    volatile float g2s = 2.0f * 2.0f;
    g2s = g2s * s;
    volatile float r = (float)((int)g2s) / 2.0f;
    r = r + 1.0f;
    volatile float tss = 2.0f * s;
    tss = tss * s;
    volatile float two_s = 2.0f * s;
    const float tso = java_two_sigma_plus_one(two_s);
    if (sigma) *sigma = s;
    if (radius) *radius = r;
    if (tSS) *tSS = tss;
    if (tSO) *tSO = tso;
    return ESRRF_OK;
}

int esrrf_bic_tso_text(float fwhm, char *buf, int len)
{
    if (!buf || len <= 0) return ESRRF_ERR_INVALID;
    volatile float s = fwhm / 2.354f;
    volatile float two_s = 2.0f * s;
    const std::string t = java_float_to_string(two_s) + "1";
    if ((int)t.size() >= len) return ESRRF_ERR_INVALID;
    snprintf(buf, (size_t)len, "%s", t.c_str());
    return ESRRF_OK;
}

int esrrf_fht_constants(float fwhm, int magnification, float *sigma, float *radius, float *tSS,
                        float *tSO)
{
    // LiveSRRF_CL.java:491-500
    const float Mf = (float)magnification;
    volatile float s = fwhm / 2.354f;
    volatile float t2 = 2.0f * s;
    volatile float ar = (float)((int)t2);
    ar = ar + 1.0f;
    volatile float tss = 2.0f * s;
    tss = tss * Mf;
    tss = tss * s;
    tss = tss * Mf;
    volatile float tso = 2.0f * s;
    tso = tso + 1.0f;
    tso = tso * Mf;
    volatile float r = Mf * ar;
    if (sigma) *sigma = s;
    if (radius) *radius = r;
    if (tSS) *tSS = tss;
    if (tSO) *tSO = tso;
    return ESRRF_OK;
}

int esrrf_plan_rolling(int n_slices, int n_frame_for_srrf_user, int frame_gap_user, int rolling,
                       int max_frames_on_gpu, int *out6)
{
    if (!out6 || n_slices < 1 || max_frames_on_gpu < 1 || n_frame_for_srrf_user < 0 ||
        frame_gap_user < 0)
        return ESRRF_ERR_INVALID;
    // gui/LiveSRRF_optimised_.java:621-627
    const int nF = n_frame_for_srrf_user == 0 ? n_slices : std::min(n_slices, n_frame_for_srrf_user);
    const int gap = (frame_gap_user == 0 || !rolling) ? nF : frame_gap_user;
    volatile float q = (float)(n_slices - nF) / (float)gap;
    const int nSR = (int)q + 1;
    // :856-858 frames per load from the memory-limited maximum
    volatile float a1 = (float)nF / (float)max_frames_on_gpu;
    int nOn = (int)std::ceil((double)a1);
    volatile float a2 = (float)nF / (float)nOn;
    nOn = (int)std::ceil((double)a2);
    nOn = std::min(nF, nOn);
    // :644 loads per SR frame
    volatile float a3 = (float)nF / (float)nOn;
    const int nLoads = (int)std::ceil((double)a3);
    out6[0] = nF;
    out6[1] = gap;
    out6[2] = nSR;
    out6[3] = nOn;
    out6[4] = nLoads;
    out6[5] = 0;
    return ESRRF_OK;
}

int esrrf_plan_load(const int *plan6, int r, int l, int *first_index, int *n_frames)
{
    if (!plan6 || r < 1 || r > plan6[2] || l < 0 || l >= plan6[4]) return ESRRF_ERR_INVALID;
    // indexStartSRRFframe = (r-1)*gap + 1; nFrameToLoad = min(nOn, nF - nOn*l)  (:281,311)
    const int start = (r - 1) * plan6[1] + 1;
    if (first_index) *first_index = start + plan6[3] * l;
    if (n_frames) *n_frames = std::min(plan6[3], plan6[0] - plan6[3] * l);
    return ESRRF_OK;
}

}  // extern "C"
