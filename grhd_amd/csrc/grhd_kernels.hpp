// grhd_kernels.hpp -- launch interface of the device kernels (grhd_kernels.hip).
#pragma once
#include "grhd_physics.hpp"

namespace grhd {

template <int MK>
void launch_boundaries_mk(const Params &P, hipStream_t s, double *p, const double *injector);

// The stage kernels (face fluxes + cell update) with FP64 or FP32 faces.
void launch_stage_f64(const Params &P, hipStream_t s, double *p, void *const F[3],
                      const int64_t fstr[3], const double *u0, double *u, int stage, double dt,
                      unsigned long long *counters, bool fused);
void launch_stage_f32(const Params &P, hipStream_t s, double *p, void *const F[3],
                      const int64_t fstr[3], const double *u0, double *u, int stage, double dt,
                      unsigned long long *counters, bool fused);

// Grids of at most this many cells use the fused stage form (one face launch,
// one cell launch); larger ones one launch per face direction and separate
// update and con2prim kernels.
constexpr int64_t kFusedStageCells = 1 << 16;

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
