// mode3d_internal.h -- entry points of the 3-D (multi-focus) pipeline.
#pragma once
#include "modes.h"

namespace esrrf {

int mode3d_create(const ModeCreate &c, ModeState &ms, float **d_px, float2 **d_U);
int mode3d_prepare(const ModeCtx &c, ModeState &ms, const void *src, int dtype, int n);
int mode3d_rgc(const ModeCtx &c, ModeState &ms, const ModeRgc &r);
int mode3d_read_gradients(const ModeCtx &c, ModeState &ms, float *out);

}  // namespace esrrf
