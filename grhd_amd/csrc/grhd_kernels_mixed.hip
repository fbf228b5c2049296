// grhd_kernels_mixed.hip -- the RK stage with FP32 reconstruction and Riemann
// solver (face fluxes stored as FP32); the cell update, sources, con2prim and
// the RK accumulation stay FP64.  Compiled with FMA contraction: this path is
// held to an error bound against the FP64 path, not to bitwise agreement.
#include "grhd_stage.hpp"

namespace grhd {

void launch_stage_f32(const Params &P, hipStream_t s, double *p, void *const F[3],
                      const int64_t fstr[3], const double *u0, double *u, int stage, double dt,
                      unsigned long long *counters, bool fused)
{
    GRHD_STAGE_DISPATCH(float)
}

void launch_face_dir_f32(const Params &P, hipStream_t s, const double *p, void *const F[3], const int64_t fstr[3],
                         int dir)
{
    GRHD_FACE_DIR_DISPATCH(float)
}

}  // namespace grhd
