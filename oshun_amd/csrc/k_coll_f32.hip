// Collision kernels, FP32 storage (FP64 arithmetic).
// Explicit instantiations of the kernel templates listed in instances.hpp.
#define OSHUN_FAMILY_TU
#include "instances.hpp"

OSHUN_FAMILY_COLL(, float)
