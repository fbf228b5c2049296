// grhd_stage_dispatch.hip -- selects the stage and face launchers of a
// problem's metric kind; the launchers themselves are instantiated by the
// per-kind units grhd_stage_{f64,f32}_{mink,ef,ks,kslog}.hip.
#include "grhd_kernels.hpp"

namespace grhd {

namespace {

template <typename R>
void stage_of_kind(const Params &P, hipStream_t s, double *p, void *const F[3], const int64_t fstr[3],
                   const double *u0, double *u, int stage, double dt, unsigned long long *counters, bool fused)
{
    switch (P.mk) {
    case MK_MINK_CART: launch_stage_mk<MK_MINK_CART, R>(P, s, p, F, fstr, u0, u, stage, dt, counters, fused); break;
    case MK_MINK_CYL: launch_stage_mk<MK_MINK_CYL, R>(P, s, p, F, fstr, u0, u, stage, dt, counters, fused); break;
    case MK_MINK_SPH: launch_stage_mk<MK_MINK_SPH, R>(P, s, p, F, fstr, u0, u, stage, dt, counters, fused); break;
    case MK_MINK_SPH_LOG:
        launch_stage_mk<MK_MINK_SPH_LOG, R>(P, s, p, F, fstr, u0, u, stage, dt, counters, fused);
        break;
    case MK_EF: launch_stage_mk<MK_EF, R>(P, s, p, F, fstr, u0, u, stage, dt, counters, fused); break;
    case MK_EF_LOG: launch_stage_mk<MK_EF_LOG, R>(P, s, p, F, fstr, u0, u, stage, dt, counters, fused); break;
    case MK_KS: launch_stage_mk<MK_KS, R>(P, s, p, F, fstr, u0, u, stage, dt, counters, fused); break;
    default: launch_stage_mk<MK_KS_LOG, R>(P, s, p, F, fstr, u0, u, stage, dt, counters, fused); break;
    }
}

template <typename R>
void face_of_kind(const Params &P, hipStream_t s, const double *p, void *const F[3], const int64_t fstr[3], int dir)
{
    switch (P.mk) {
    case MK_MINK_CART: launch_face_dir_mk<MK_MINK_CART, R>(P, s, p, F, fstr, dir); break;
    case MK_MINK_CYL: launch_face_dir_mk<MK_MINK_CYL, R>(P, s, p, F, fstr, dir); break;
    case MK_MINK_SPH: launch_face_dir_mk<MK_MINK_SPH, R>(P, s, p, F, fstr, dir); break;
    case MK_MINK_SPH_LOG: launch_face_dir_mk<MK_MINK_SPH_LOG, R>(P, s, p, F, fstr, dir); break;
    case MK_EF: launch_face_dir_mk<MK_EF, R>(P, s, p, F, fstr, dir); break;
    case MK_EF_LOG: launch_face_dir_mk<MK_EF_LOG, R>(P, s, p, F, fstr, dir); break;
    case MK_KS: launch_face_dir_mk<MK_KS, R>(P, s, p, F, fstr, dir); break;
    default: launch_face_dir_mk<MK_KS_LOG, R>(P, s, p, F, fstr, dir); break;
    }
}

template <typename R>
void zstage_of_kind(const Params &P, hipStream_t s, const double *pin, double *pout, void *const F[3],
                    const int64_t fstr[3], void *geo_z, const double *u0, double *u, int stage, double dt,
                    unsigned long long *counters)
{
    switch (P.mk) {
    case MK_MINK_CART: launch_stage_z_mk<MK_MINK_CART, R>(P, s, pin, pout, F, fstr, geo_z, u0, u, stage, dt, counters); break;
    case MK_MINK_CYL: launch_stage_z_mk<MK_MINK_CYL, R>(P, s, pin, pout, F, fstr, geo_z, u0, u, stage, dt, counters); break;
    case MK_MINK_SPH: launch_stage_z_mk<MK_MINK_SPH, R>(P, s, pin, pout, F, fstr, geo_z, u0, u, stage, dt, counters); break;
    case MK_MINK_SPH_LOG:
        launch_stage_z_mk<MK_MINK_SPH_LOG, R>(P, s, pin, pout, F, fstr, geo_z, u0, u, stage, dt, counters);
        break;
    case MK_EF: launch_stage_z_mk<MK_EF, R>(P, s, pin, pout, F, fstr, geo_z, u0, u, stage, dt, counters); break;
    case MK_EF_LOG: launch_stage_z_mk<MK_EF_LOG, R>(P, s, pin, pout, F, fstr, geo_z, u0, u, stage, dt, counters); break;
    case MK_KS: launch_stage_z_mk<MK_KS, R>(P, s, pin, pout, F, fstr, geo_z, u0, u, stage, dt, counters); break;
    default: launch_stage_z_mk<MK_KS_LOG, R>(P, s, pin, pout, F, fstr, geo_z, u0, u, stage, dt, counters); break;
    }
}

}  // namespace

void launch_stage_z(const Params &P, hipStream_t s, const double *pin, double *pout, void *const F[3],
                    const int64_t fstr[3], void *geo_z, const double *u0, double *u, int stage, double dt,
                    unsigned long long *counters, bool mixed)
{
    if (mixed) zstage_of_kind<float>(P, s, pin, pout, F, fstr, geo_z, u0, u, stage, dt, counters);
    else zstage_of_kind<double>(P, s, pin, pout, F, fstr, geo_z, u0, u, stage, dt, counters);
}

void launch_stage_f64(const Params &P, hipStream_t s, double *p, void *const F[3], const int64_t fstr[3],
                      const double *u0, double *u, int stage, double dt, unsigned long long *counters, bool fused)
{
    stage_of_kind<double>(P, s, p, F, fstr, u0, u, stage, dt, counters, fused);
}

void launch_stage_f32(const Params &P, hipStream_t s, double *p, void *const F[3], const int64_t fstr[3],
                      const double *u0, double *u, int stage, double dt, unsigned long long *counters, bool fused)
{
    stage_of_kind<float>(P, s, p, F, fstr, u0, u, stage, dt, counters, fused);
}

void launch_face_dir_f64(const Params &P, hipStream_t s, const double *p, void *const F[3], const int64_t fstr[3],
                         int dir)
{
    face_of_kind<double>(P, s, p, F, fstr, dir);
}

void launch_face_dir_f32(const Params &P, hipStream_t s, const double *p, void *const F[3], const int64_t fstr[3],
                         int dir)
{
    face_of_kind<float>(P, s, p, F, fstr, dir);
}

}  // namespace grhd
