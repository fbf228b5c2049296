// grhd_batch_f64_b.hip -- batched stage kernels with FP64 faces, Eddington-Finkelstein and Kerr-Schild kinds.
#include "grhd_batch_kernels.hpp"

namespace grhd {
GRHD_BATCH_DISPATCH(double, b, MK_EF, MK_EF_LOG, MK_KS, MK_KS_LOG)
}  // namespace grhd
