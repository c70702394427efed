// mode3d.cpp -- host side of the 3-D (multi-focus) pipeline.
#include "mode3d_internal.h"

namespace esrrf {

int mode3d_create(const ModeCreate &, ModeState &ms, float **, float2 **)
{
    ms.err = "3-D mode is not available in this build";
    return ESRRF_ERR_UNSUPPORTED;
}
int mode3d_prepare(const ModeCtx &, ModeState &ms, const void *, int, int)
{
    ms.err = "3-D mode is not available in this build";
    return ESRRF_ERR_UNSUPPORTED;
}
int mode3d_rgc(const ModeCtx &, ModeState &ms, const ModeRgc &)
{
    ms.err = "3-D mode is not available in this build";
    return ESRRF_ERR_UNSUPPORTED;
}
int mode3d_read_gradients(const ModeCtx &, ModeState &ms, float *)
{
    ms.err = "3-D mode is not available in this build";
    return ESRRF_ERR_UNSUPPORTED;
}

}  // namespace esrrf
