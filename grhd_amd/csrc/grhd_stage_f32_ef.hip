// grhd_stage_f32_ef.hip -- RK stage kernels with FP32 faces (mixed precision; compiled with contraction): Eddington-Finkelstein.
#include "grhd_stage.hpp"

namespace grhd {
GRHD_STAGE_UNIT(MK_EF, float)
GRHD_STAGE_UNIT(MK_EF_LOG, float)
}  // namespace grhd
