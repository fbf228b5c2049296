// grhd_batch_mixed_a.hip -- batched stage kernels with FP32 faces (mixed precision), Minkowski metric kinds.
#include "grhd_batch_kernels.hpp"

namespace grhd {
GRHD_BATCH_DISPATCH(float, a, MK_MINK_CART, MK_MINK_CYL, MK_MINK_SPH, MK_MINK_SPH_LOG)
}  // namespace grhd
