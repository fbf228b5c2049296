// grhd_stage_f64_ef.hip -- RK stage kernels with FP64 faces (the reference's rounding sequence): Eddington-Finkelstein.
#include "grhd_stage.hpp"

namespace grhd {
GRHD_STAGE_UNIT(MK_EF, double)
GRHD_STAGE_UNIT(MK_EF_LOG, double)
}  // namespace grhd
