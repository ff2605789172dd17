// The l-streaming stage kernel of the general system at m0 = 1: every
// coupling program in FP64, the full program in FP32.  Explicit
// instantiations of the kernel templates listed in instances.hpp.
#define OSHUN_FAMILY_TU
#include "instances.hpp"

OSHUN_LM_KINDS(, double, 1)
OSHUN_LM_TD(, float, 1, 0)
