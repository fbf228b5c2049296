// grhd_stage_f32_ks.hip -- RK stage kernels with FP32 faces (mixed precision; compiled with contraction): Kerr-Schild, physical r.
#include "grhd_stage.hpp"

namespace grhd {
GRHD_STAGE_UNIT(MK_KS, float)
}  // namespace grhd
