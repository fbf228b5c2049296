// grhd_stage_f64_ks.hip -- RK stage kernels with FP64 faces (the reference's rounding sequence): Kerr-Schild, physical r.
#include "grhd_stage.hpp"

namespace grhd {
GRHD_STAGE_UNIT(MK_KS, double)
}  // namespace grhd
