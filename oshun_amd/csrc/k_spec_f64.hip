// General-system stage kernels (interpreter and specialised shapes), FP64.
// Explicit instantiations of the kernel templates listed in instances.hpp.
#define OSHUN_FAMILY_TU
#include "instances.hpp"

#define OSHUN_SPEC_EXT
OSHUN_STAGE_SPECS(OSHUN_SPEC_F64)
OSHUN_I_GEN(, double)
