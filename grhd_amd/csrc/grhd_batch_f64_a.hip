// grhd_batch_f64_a.hip -- batched stage kernels with FP64 faces, Minkowski metric kinds.
#include "grhd_batch_kernels.hpp"

namespace grhd {
GRHD_BATCH_DISPATCH(double, a, MK_MINK_CART, MK_MINK_CYL, MK_MINK_SPH, MK_MINK_SPH_LOG)
}  // namespace grhd
