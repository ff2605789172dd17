// Single-run stage pairs marched per wavefront (vlasov_march_1d), storage type float.
// Explicit instantiations of the kernel templates listed in instances.hpp.
#define OSHUN_FAMILY_TU
#include "instances.hpp"

OSHUN_FAMILY_MARCH(, float)
