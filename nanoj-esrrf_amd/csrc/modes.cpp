// modes.cpp -- dispatch of the FHT and 3-D pipelines (filled in by k_fht.hip / k_3d.hip).
#include "modes.h"

namespace esrrf {

int mode_create(const ModeCreate &c, ModeState &ms, float **, float2 **, float **, size_t *)
{
    ms.mode = c.mode;
    ms.err = "mode not implemented yet";
    return ESRRF_ERR_UNSUPPORTED;
}
int mode_prepare(int, const ModeCtx &, ModeState &ms, const void *, int, int)
{
    ms.err = "mode not implemented yet";
    return ESRRF_ERR_UNSUPPORTED;
}
int mode_rgc(int, const ModeCtx &, ModeState &ms, const ModeRgc &)
{
    ms.err = "mode not implemented yet";
    return ESRRF_ERR_UNSUPPORTED;
}
int mode_read_gradients(int, const ModeCtx &, ModeState &ms, float *)
{
    ms.err = "mode not implemented yet";
    return ESRRF_ERR_UNSUPPORTED;
}
long long mode_active_samples(int, const ModeState &ms) { return ms.active; }
void free_mode_state(ModeState &) {}

}  // namespace esrrf
