// grhd_stage_f64_mink.hip -- RK stage kernels with FP64 faces (the reference's rounding sequence): the Minkowski kinds.
#include "grhd_stage.hpp"

namespace grhd {
GRHD_STAGE_UNIT(MK_MINK_CART, double)
GRHD_STAGE_UNIT(MK_MINK_CYL, double)
GRHD_STAGE_UNIT(MK_MINK_SPH, double)
GRHD_STAGE_UNIT(MK_MINK_SPH_LOG, double)
}  // namespace grhd
