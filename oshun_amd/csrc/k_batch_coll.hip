// Batched collision kernels.
// Explicit instantiations of the kernel templates listed in instances.hpp.
#define OSHUN_FAMILY_TU
#include "instances.hpp"

OSHUN_FAMILY_COLLB(, double)
OSHUN_FAMILY_COLLB(, float)
