// grhd_stage_f64_kslog.hip -- RK stage kernels with FP64 faces (the reference's rounding sequence): Kerr-Schild, logarithmic r.
#include "grhd_stage.hpp"

namespace grhd {
GRHD_STAGE_UNIT(MK_KS_LOG, double)
}  // namespace grhd
