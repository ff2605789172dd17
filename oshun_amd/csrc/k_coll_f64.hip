// Collision kernels, FP64 storage.
// Explicit instantiations of the kernel templates listed in instances.hpp.
#define OSHUN_FAMILY_TU
#include "instances.hpp"

OSHUN_FAMILY_COLL(, double)
