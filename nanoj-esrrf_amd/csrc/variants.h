// variants.h -- alternative kernels kept beside the defaults so that tests can show they give
// the same results bit for bit (per-point gradient stencils, scattered FHT stores, per-voxel
// 3-D kernels, ...).  Selected by ESRRF_* environment variables that are read once, when an
// engine (or sweep) is created; launches only see the copy it holds.
#pragma once

namespace esrrf {

struct KernelVariants {
    int bic_prep = 4;             // ESRRF_BIC_PREP=1: one 4x4 stencil per sample point;
                                  // 2: shared row sums, one frame per block
    int bic_stage = 0;            // ESRRF_BIC_STAGE=1: single-frame kernel stages in registers
    int bic_pair_direct = 0;      // ESRRF_BIC_PAIR_DIRECT=1: pair kernel fills after compute
    int fht_scatter = 0;          // ESRRF_FHT_SCATTER=1: pixel-by-pixel last FHT pass
    int fht_xcd = 1;              // ESRRF_FHT_XCD=0: column passes in launch order
    int fht_prefetch = 1;         // ESRRF_FHT_PREFETCH=0: no register prefetch, last pass
    int fht_pair = 1;             // ESRRF_FHT_PAIR=0: single-pixel FHT RGC kernel
    int k3d_kernel = 3;           // ESRRF_3D_KERNEL=1/2: per-voxel 3-D RGC kernels
    int k3d_split = 2;            // ESRRF_3D_SPLIT=1: one wave per phase in k3d_rgc3
    int k3d_zpair = 1;            // ESRRF_3D_ZPAIR=0: one z-slice per tiled 3-D block
    int sweep_per_fwhm = 0;       // ESRRF_SWEEP_PER_FWHM=1: one sweep emit launch per fwhm
};

// The variants the environment selects (defaults when unset).
KernelVariants kernel_variants_from_env();

}  // namespace esrrf
