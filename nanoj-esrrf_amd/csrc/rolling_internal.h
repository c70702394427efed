// rolling_internal.h -- launch parameters of the rolling-window accumulation kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace esrrf {

constexpr int ROLL_GROUP = 16;    // windows per launch, held in registers per pixel
constexpr int ROLL_MAX_WIN = 64;  // windows one chunk may touch (launched in groups)

// The launch covers the consecutive windows r0 .. r0 + nwin - 1; window r starts at frame
// r gap of the stack, the chunk at frame c0, and window r keeps its cumulants in slot
// r mod nslots of `state` ([nslots][5][M*M][h][w]: AVG, E2, TAC2, INT, Prev, in the emit
// order of the frame values) until it is complete.
struct RollAccArgs {
    const float *C, *V;           // emitted CGLH / intensity of the chunk [nframes][M*M][h][w]
    float *state;
    int w, h, M;
    int nframes;                  // frames of the chunk
    int nF;                       // frames per SR window
    int r0, nwin;                 // first window, windows (<= ROLL_GROUP)
    int c0, gap, nslots;
    float sens;
    int sens_int;
    int int_weighting;
    float rn, rn1;                // 1/nF, 1/(nF - 1)
};

hipError_t launch_roll_acc(const RollAccArgs &a, hipStream_t s);
// channels 0..3 of a complete window's slot to the image order [4][hM][wM] of dst
hipError_t launch_roll_to_image(const float *src, float *dst, int w, int h, int M, hipStream_t s);

}  // namespace esrrf
