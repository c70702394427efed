// rolling_internal.h -- launch parameters of the rolling-window accumulation kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace esrrf {

constexpr int ROLL_GROUP = 8;     // windows per thread
constexpr int ROLL_MAX_WIN = 64;  // windows per launch

struct RollAccArgs {
    const float *C, *V;           // emitted CGLH / intensity of the chunk [nframes][M*M][h][w]
    int w, h, M;
    int nframes;                  // frames of the chunk
    int nF;                       // frames per SR window
    int nwin;                     // windows in this launch
    int k0[ROLL_MAX_WIN];         // index inside each window of the chunk's first frame
    float *state[ROLL_MAX_WIN];   // each window's [5][hM][wM]: AVG, E2, TAC2, INT, Prev
    float sens;
    int sens_int;
    int int_weighting;
    float rn, rn1;                // 1/nF, 1/(nF - 1)
};

hipError_t launch_roll_acc(const RollAccArgs &a, hipStream_t s);

}  // namespace esrrf
