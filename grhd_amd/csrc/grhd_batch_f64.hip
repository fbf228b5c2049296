// grhd_batch_f64.hip -- batched stage kernels with FP64 faces (reference
// rounding sequence: compiled without FMA contraction).
#include "grhd_batch_kernels.hpp"

namespace grhd {
GRHD_BATCH_DISPATCH(double)
}  // namespace grhd
