// Batched stage-pair kernels.
// Explicit instantiations of the kernel templates listed in instances.hpp.
#define OSHUN_FAMILY_TU
#include "instances.hpp"

OSHUN_FAMILY_PAIRB(, double)
OSHUN_FAMILY_PAIRB(, float)
OSHUN_FAMILY_PAIRB_F32_R4()
