// variants.cpp -- the one place the ESRRF_* kernel-variant switches are read.
#include "variants.h"

#include <cstdlib>

namespace esrrf {

namespace {

int env_int(const char *name, int dflt)
{
    const char *v = std::getenv(name);
    return v ? std::atoi(v) : dflt;
}

}  // namespace

KernelVariants kernel_variants_from_env()
{
    KernelVariants k;
    k.bic_prep = env_int("ESRRF_BIC_PREP", 4);
    if (k.bic_prep != 1 && k.bic_prep != 2) k.bic_prep = 4;
    k.bic_stage = env_int("ESRRF_BIC_STAGE", 0) != 0;
    k.bic_pair_direct = env_int("ESRRF_BIC_PAIR_DIRECT", 0) != 0;
    k.fht_scatter = env_int("ESRRF_FHT_SCATTER", 0) != 0;
    k.fht_xcd = env_int("ESRRF_FHT_XCD", 1) != 0;
    k.fht_prefetch = env_int("ESRRF_FHT_PREFETCH", 1) != 0;
    k.fht_pair = env_int("ESRRF_FHT_PAIR", 1) != 0;
    k.k3d_kernel = env_int("ESRRF_3D_KERNEL", 3);
    k.k3d_split = env_int("ESRRF_3D_SPLIT", 2) == 1 ? 1 : 2;
    k.k3d_zpair = env_int("ESRRF_3D_ZPAIR", 1) != 0;
    k.sweep_per_fwhm = env_int("ESRRF_SWEEP_PER_FWHM", 0) == 1;
    return k;
}

}  // namespace esrrf
