// grhd_kernels.hpp -- launch interface of the device kernels (grhd_kernels.hip).
#pragma once
#include "grhd_physics.hpp"

namespace grhd {

template <int MK>
void launch_boundaries_mk(const Params &P, hipStream_t s, double *p, const double *injector);

// The stage launcher and the single-direction face launcher of one metric kind
// and face arithmetic R (double: FP64 faces; float: FP32 faces), defined in
// grhd_stage.hpp and instantiated per kind by the grhd_stage_*.hip units.
template <int MK, typename R>
void launch_stage_mk(const Params &P, hipStream_t s, double *p, void *const F[3], const int64_t fstr[3],
                     const double *u0, double *u, int stage, double dt, unsigned long long *counters,
                     bool fused);
template <int MK, typename R>
void launch_face_dir_mk(const Params &P, hipStream_t s, const double *p, void *const F[3], const int64_t fstr[3],
                        int dir);
template <int MK, typename R>
void launch_stage_z_mk(const Params &P, hipStream_t s, const double *pin, double *pout, void *const F[3],
                       const int64_t fstr[3], void *geo_z, const double *u0, double *u, int stage, double dt,
                       unsigned long long *counters);

// The stage kernels (face fluxes + cell update) with FP64 or FP32 faces.
void launch_stage_f64(const Params &P, hipStream_t s, double *p, void *const F[3],
                      const int64_t fstr[3], const double *u0, double *u, int stage, double dt,
                      unsigned long long *counters, bool fused);
void launch_stage_f32(const Params &P, hipStream_t s, double *p, void *const F[3],
                      const int64_t fstr[3], const double *u0, double *u, int stage, double dt,
                      unsigned long long *counters, bool fused);

// Grids of at most this many cells use the fused stage form (one face launch,
// one cell launch); larger ones one launch per face direction and separate
// update and con2prim kernels.  The batch also splits its groups there.
constexpr int64_t kFusedStageCells = 1 << 16;
// A single problem with an inactive y or z axis keeps the fused form up to
// this many cells: measured faster on 2D equatorial tori of 80k, 320k and
// 1.28M cells (+26, +19 and +15 % in FP64), while 3D tori above 2^16 cells
// run faster with one launch per direction (the fused form lost 1 % at 524k
// cells, 8 % at 1.77M and 12 % at 4.2M).
constexpr int64_t kFusedStage2DCells = 1 << 21;

inline bool single_stage_fused(const Params &P)
{
    const int64_t n = (int64_t)P.nx * P.ny * P.nz;
    return n <= kFusedStageCells || ((P.ny == 1 || P.nz == 1) && n <= kFusedStage2DCells);
}

// The single-problem face kernels can defer their HLLC fallbacks to a second
// kernel (Params::defer_list): on grids with face-geometry tables, for the
// compiled MP5 + sensor + HLLC scheme on the Kerr-Schild metrics, with face
// indices that fit 32 bits.
inline bool defer_faces(const Params &P)
{
    const int64_t nmax = (int64_t)(P.nx + 1) * (P.ny + 1) * (P.nz + 1);
    return !single_stage_fused(P) && (P.mk == MK_KS || P.mk == MK_KS_LOG) && P.rec == 4 && P.sensor != 0 &&
           P.rs == 2 && nmax < ((int64_t)1 << 32);
}

// The face kernel of direction dir (1..3) alone (face-tiling tuner).
void launch_face_dir_f64(const Params &P, hipStream_t s, const double *p, void *const F[3], const int64_t fstr[3],
                         int dir);
void launch_face_dir_f32(const Params &P, hipStream_t s, const double *p, void *const F[3], const int64_t fstr[3],
                         int dir);

// One RK stage (1..3) of a 3D grid in the z-march form (P.zseg > 0): reads
// the primitives from pin and writes them to pout (interior only; the ghost
// fill of pout follows separately).  geo_z: the z-face geometry table.
void launch_stage_z(const Params &P, hipStream_t s, const double *pin, double *pout, void *const F[3],
                    const int64_t fstr[3], void *geo_z, const double *u0, double *u, int stage, double dt,
                    unsigned long long *counters, bool mixed);

// One RK stage: face fluxes, cell update (stage 1..3; 0 = rhs only into u),
// then the ghost fill when with_boundaries.
void launch_stage(const Params &P, hipStream_t s, double *p, void *const F[3], const int64_t fstr[3],
                  const double *u0, double *u, int stage, double dt, unsigned long long *counters,
                  bool mixed, const double *injector, bool with_boundaries);
void launch_boundaries(const Params &P, hipStream_t s, double *p, const double *injector);
void launch_p2c(const Params &P, hipStream_t s, const double *p, double *u);
void launch_c2p(const Params &P, hipStream_t s, double *p, double *u, unsigned long long *counters);
// Returns the number of partial records (4 doubles each) written.
int launch_cfl(const Params &P, hipStream_t s, double *partial);
// block partial sums of the mass monitor (at most max_blocks); returns the block count
int launch_mass(const Params &P, hipStream_t s, const double *D, double *partial, int max_blocks);
// observers: global diagnostics (10 partials per block), per-plane masses (one
// per phi plane), accretion-rate partials on the probe cell
constexpr int kGlobalPartials = 10;
constexpr int kPpiModes = 4;  // PPI_MAX_MODE (variables.f90:141)
int launch_global(const Params &P, hipStream_t s, const double *p, const double *u, double *partial, int max_blocks);
void launch_plane_mass(const Params &P, hipStream_t s, const double *p, const double *u, double *partial);
int launch_mdot(const Params &P, hipStream_t s, int probe, const double *p, const double *u, double *partial,
                int max_blocks);

}  // namespace grhd
