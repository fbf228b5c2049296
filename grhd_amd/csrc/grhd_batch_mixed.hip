// grhd_batch_mixed.hip -- batched stage kernels with FP32 faces (mixed
// precision: compiled with FMA contraction, like grhd_kernels_mixed.hip).
#include "grhd_batch_kernels.hpp"

namespace grhd {
GRHD_BATCH_DISPATCH(float)
}  // namespace grhd
