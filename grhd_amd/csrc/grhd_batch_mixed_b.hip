// grhd_batch_mixed_b.hip -- batched stage kernels with FP32 faces (mixed precision), Eddington-Finkelstein and Kerr-Schild kinds.
#include "grhd_batch_kernels.hpp"

namespace grhd {
GRHD_BATCH_DISPATCH(float, b, MK_EF, MK_EF_LOG, MK_KS, MK_KS_LOG)
}  // namespace grhd
