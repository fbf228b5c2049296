// grhd_stage_f32_mink.hip -- RK stage kernels with FP32 faces (mixed precision; compiled with contraction): the Minkowski kinds.
#include "grhd_stage.hpp"

namespace grhd {
GRHD_STAGE_UNIT(MK_MINK_CART, float)
GRHD_STAGE_UNIT(MK_MINK_CYL, float)
GRHD_STAGE_UNIT(MK_MINK_SPH, float)
GRHD_STAGE_UNIT(MK_MINK_SPH_LOG, float)
}  // namespace grhd
