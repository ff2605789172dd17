// Single-run 1D stage and stage-pair kernels, FP64 (with the implicit-E modes).
// Explicit instantiations of the kernel templates listed in instances.hpp.
#define OSHUN_FAMILY_TU
#include "instances.hpp"

OSHUN_FAMILY_PAIR(, double)
OSHUN_FAMILY_PAIR_IMP()
