// Batched marched stage pairs (vlasov_march_1d_batch), FP64 and FP32 storage.
// Explicit instantiations of the kernel templates listed in instances.hpp.
#define OSHUN_FAMILY_TU
#include "instances.hpp"

OSHUN_FAMILY_MARCHB(, double)
OSHUN_FAMILY_MARCHB(, float)
