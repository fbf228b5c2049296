// grhd_stage_f32_kslog.hip -- RK stage kernels with FP32 faces (mixed precision; compiled with contraction): Kerr-Schild, logarithmic r.
#include "grhd_stage.hpp"

namespace grhd {
GRHD_STAGE_UNIT(MK_KS_LOG, float)
}  // namespace grhd
