// device_tiles.h -- XCD-aware order of the tiles of a one-dimensional grid of tile blocks.
#pragma once
#include <hip/hip_runtime.h>

namespace esrrf {

// Tile index of this block among n tiles (row-major).  The chip hands workgroups to its
// eight XCDs in turn (workgroup i to XCD i mod 8), each XCD with its own L2, while
// neighbouring tiles share the halos of their staged inputs.  Where n divides by eight, XCD j
// takes the contiguous tile range [j n/8, (j+1) n/8) in order instead, so the blocks resident
// on one XCD cover neighbouring tiles and fetch the shared halos through one L2.  Every tile
// is visited once in either order.
__device__ __forceinline__ int xcd_tile_1d(int n)
{
    const int b = (int)blockIdx.x;
    return (n & 7) == 0 ? (b & 7) * (n >> 3) + (b >> 3) : b;
}

}  // namespace esrrf
