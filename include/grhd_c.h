/*
 * grhd_c.h -- C ABI of the MI355X GRHD2 finite-volume update (libgrhd_hip.so).
 *
 * The reference has no FFI: its update path is a set of Fortran module
 * procedures sharing module state.  Each entry point below replaces one of
 * those procedures (or one block of the reference main program) and is what a
 * Fortran driver binds through iso_c_binding (grhd_amd/fortran/grhd_c.f90)
 * and what the Python tests bind through ctypes.
 *
 * Conventions
 *  - Every function returns an int status: GRHD_OK (0) or a GRHD_ERR_* code.
 *    No exception crosses the ABI.  grhd_last_error() describes the last
 *    failure of a context (or of grhd_create when ctx is NULL).
 *  - A context owns all device memory and one HIP stream.  Host buffers are
 *    borrowed for the duration of a call only.
 *  - Host arrays use the reference's own Fortran layout:
 *      primitives p(neq, -ng:nx+ng, -ng:ny+ng, -ng:nz+ng)   (neq = 5 fastest)
 *      conserved  u(neq, nx, ny, nz)
 *    i.e. C arrays [nz+2ng+1][ny+2ng+1][nx+2ng+1][5] and [nz][ny][nx][5].
 *    Variables: 0 rho / D, 1 p / tau, 2..4 v^i / S_i (variables.f90:24-30).
 *  - A context is used by one host thread at a time; distinct contexts are
 *    independent.
 */
#ifndef GRHD_C_H
#define GRHD_C_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define GRHD_OK 0
#define GRHD_ERR_INVALID_CONFIG 1   /* what parameters.f90:359-588 rejects */
#define GRHD_ERR_INVALID_METRIC 2   /* evolution.f90:211-272 error stops */
#define GRHD_ERR_OUT_OF_MEMORY 3
#define GRHD_ERR_HIP 4
#define GRHD_ERR_INVALID_ARGUMENT 5
#define GRHD_ERR_UNSUPPORTED 6
#define GRHD_ERR_CAPACITY 7         /* batch: no free slot or HBM budget */

/* identifiers: same values as the reference (variables.f90:224-237) */
#define GRHD_REC_GODUNOV 1
#define GRHD_REC_TVD 2
#define GRHD_REC_WENO3 3
#define GRHD_REC_MP5 4
#define GRHD_REC_WENO5 5
#define GRHD_RS_HLLE 1
#define GRHD_RS_HLLC 2
#define GRHD_LIM_MINMOD 1
#define GRHD_LIM_SUPERBEE 2
#define GRHD_LIM_MC 3
#define GRHD_METRIC_MINKOWSKI 0
#define GRHD_METRIC_EDDINGTON_FINKELSTEIN 1
#define GRHD_METRIC_KERR_SCHILD 2
#define GRHD_GEOM_CARTESIAN 0
#define GRHD_GEOM_CYLINDRICAL 1
#define GRHD_GEOM_SPHERICAL 2
#define GRHD_GEOM_SPHEROIDAL 3
/* arithmetic of the face computation (reconstruction + Riemann solver);
 * con2prim, the RK3 accumulation, sources and metric determinants are FP64
 * in both modes */
#define GRHD_PRECISION_FP64 0
#define GRHD_PRECISION_MIXED 1

/* Problem description: the fields of the .par surface the update path reads
 * (parameters.f90:146-353) after the problem preset has been applied.
 * problem: 1 sod, 2 strong_shock, 3 shu_osher, 4 kelvin_helmholtz,
 * 5 axisymmetric_jet, 6 convergence, 7 michel, 8 dust, 9 offaxis_blast,
 * 10 fishbone_equatorial, 11 fishbone_sagittal, 12 fishbone_3d
 * (parameters.f90:893-921).  Angles y_min..z_max are in radians. */
typedef struct grhd_config {
    int32_t problem;
    int32_t metric;
    int32_t geometry;
    int32_t log_r;
    double adiabatic_index;
    double bh_mass;
    double spin;
    int32_t nx, ny, nz;
    int32_t reconstruction;
    double r_min, r_max, y_min, y_max, z_min, z_max;
    int32_t riemann_solver;
    int32_t tvd_limiter;
    int32_t shock_sensor;
    int32_t precision;
    double cfl;
    double final_time;
    /* initial-condition parameters (variables.f90:166-219) */
    double sod_interface, sod_density_left, sod_density_right;
    double sod_pressure_left, sod_pressure_right;
    double strong_interface, strong_density_left, strong_density_right;
    double strong_pressure_left, strong_pressure_right;
    double shu_interface, shu_density_left, shu_pressure_left, shu_density_right;
    double shu_density_amplitude, shu_wave_number, shu_pressure_right;
    double khi_half_width, khi_density_inner, khi_density_outer;
    double khi_velocity_inner, khi_velocity_outer, khi_pressure;
    double khi_perturbation_amplitude, khi_perturbation_wave_number, khi_perturbation_width;
    double jet_ambient_density, jet_ambient_pressure, jet_ambient_velocity;
    double jet_nozzle_radius, jet_nozzle_length, jet_density, jet_pressure, jet_velocity;
    double convergence_density, convergence_amplitude, convergence_velocity;
    double convergence_pressure, convergence_wave_number;
    double michel_critical_radius, michel_critical_density;
    double dust_accretion_constant;
    double offaxis_radial_center, offaxis_phi_center, offaxis_width;
    double offaxis_background_density, offaxis_background_pressure;
    double offaxis_density_amplitude, offaxis_pressure_amplitude;
    double fm_inner_radius, fm_pressure_max_radius, fm_polytropic_constant;
    /* 1 accepts fishbone_3d with theta away from the poles (a build option,
     * the reference requires theta = [0, pi], parameters.f90:577-583).  A
     * theta boundary that is not a pole gets an outflow ghost fill (copy of
     * the edge cell, rho and p floored, v^theta limited to outflow); a pole
     * keeps the trans-polar fill (conditions.f90:1032-1050). */
    int32_t allow_open_theta;
    int32_t reserved[7];
} grhd_config;

typedef struct grhd_ctx grhd_ctx;

/* Allocates the device state for one problem.  Replaces allocate_and_grid +
 * precalculate_metric_cache + choose_numerical_architecture + the floors
 * (initialization.f90:23-391, 452-457): no metric is stored, it is evaluated
 * in the kernels.  Validation follows parameters.f90:359-588. */
int grhd_create(const grhd_config *cfg, grhd_ctx **out);
/* The validation of grhd_create alone, without touching a device (the
 * reference's --check, grhd2.f90:42-55); the message is grhd_last_error(NULL). */
int grhd_validate(const grhd_config *cfg);
int grhd_destroy(grhd_ctx *ctx);
/* Message for the last failure of ctx; ctx == NULL gives the last grhd_create failure. */
const char *grhd_last_error(const grhd_ctx *ctx);
/* dims[0..3] = nx, ny, nz, nghost */
int grhd_dims(const grhd_ctx *ctx, int32_t *dims);

/* set_initial_conditions (conditions.f90:26-214) followed by prim_to_cons over
 * the interior (grhd2.f90:126-134) and the Michel injector cache
 * (conditions.f90:229-246). */
int grhd_initial_conditions(grhd_ctx *ctx);
/* White noise on the cells denser than 5 % of the maximum density,
 * followed by U = P2C(p): type 1 pressure noise
 * (inject_pressure_noise, conditions.f90:603-636), type 2 isentropic density
 * noise (inject_density_noise, conditions.f90:730-768); type 3 additive
 * noise A (2 U - 1) on v^z (no reference counterpart: the random part of the
 * Kelvin-Helmholtz seed perturbation).  The deviates come from a
 * counter-based generator keyed by (seed, cell index); the reference's
 * compiler-runtime random_number stream is not portable. */
int grhd_perturb(grhd_ctx *ctx, int32_t type, double amplitude, uint64_t seed);
/* Polytropic constant K of the Fishbone-Moncrief torus that puts rho = 1 at
 * the pressure maximum: K = (h_c - 1)(Gamma - 1)/Gamma with
 * h_c = exp(W_in - W(r_c, pi/2)) (conditions.f90:470-546). */
int grhd_fishbone_unit_density_k(const grhd_config *cfg, double *K);
/* Lapse alpha and shift beta^i at the interior cell centres of cfg's grid,
 * computed on the host (no device, no context): alpha(nx, ny, nz) and
 * beta(3, nx, ny, nz) in Fortran order, either may be NULL.  Replaces the
 * reference's alpha_c / beta_c cache (initialization.f90:283-299) that
 * write_vtk_cell_data (output.f90:402-455) writes as Lapse_alpha and
 * Shift_beta_r. */
int grhd_lapse_shift(const grhd_config *cfg, double *alpha, double *beta);
/* Uploads a full ghosted primitive array and rebuilds U = P2C(p) on the
 * device (grhd2.f90:206-218 after a perturbation). */
int grhd_set_primitives(grhd_ctx *ctx, const double *p);
/* Uploads a full ghosted primitive array without touching U. */
int grhd_put_primitives(grhd_ctx *ctx, const double *p);
/* Uploads interior conserved variables, then cons_to_prim and boundaries
 * rebuild p (restart path, grhd2.f90:103-116). */
int grhd_set_conserved(grhd_ctx *ctx, const double *u);
/* Uploads interior conserved variables only. */
int grhd_put_conserved(grhd_ctx *ctx, const double *u);
/* Downloads the ghosted primitives.  Ghost cells that no kernel reads
 * (corners and the ghost planes of axes with n == 1) are copies of the
 * neighbouring plane. */
int grhd_get_primitives(grhd_ctx *ctx, double *p);
int grhd_get_conserved(grhd_ctx *ctx, double *u);

/* update_adaptive_dt (evolution.f90:198-275): dt = CFL / sum_d max(alpha
 * sqrt(gamma^dd) + |beta^d|)/dx^d.  Returns GRHD_ERR_INVALID_METRIC where the
 * reference error-stops. */
int grhd_compute_dt(grhd_ctx *ctx, double *dt);
/* nsteps full SSP-RK3 steps with a fixed dt (grhd2.f90:232-265: boundaries,
 * three stages of calc_rhs, RK combination, cons_to_prim, boundaries). */
int grhd_step(grhd_ctx *ctx, double dt, int32_t nsteps);
/* The reference time loop (grhd2.f90:179-232) up to t_final or max_steps:
 * dt from grhd_compute_dt, clamped at the end.  *t is read and updated. */
int grhd_advance(grhd_ctx *ctx, double t_final, int32_t max_steps, double *t,
                 int32_t *steps_done, double *last_dt);
/* Mass monitor (grhd2.f90:283-301): the sum over the interior of the current
 * U(D) dx dy dz, reduced on the device (per-block partial sums added on the
 * host in block order; the reference's serial order is not kept, so the sum
 * agrees to rounding). */
int grhd_total_mass(grhd_ctx *ctx, double *mass);
/* calc_convergence_norms (output.f90:1144-1258) at integration time t: L1, L2
 * and Linf of rho against the analytic solution on row j = max(1, ny/2),
 * k = 1 -- the advected wave (problem 6: exact cell averages, cells 1..nx),
 * Michel (7: the analytic initial data) and dust (8: the stationary free fall),
 * both on cells int(0.1 nx)..int(0.95 nx).  norms[3] = L1, L2, Linf in the
 * reference's summation order; *imax (may be NULL) = cell of the Linf.
 * GRHD_ERR_UNSUPPORTED for the other problems (the reference prints nothing
 * for them). */
int grhd_convergence_norms(grhd_ctx *ctx, double t, double *norms, int32_t *imax);
/* calc_global_diagnostics (output.f90:1052-1125) of the current state,
 * reduced on the device: out[8] = total mass, disk mass (rho > 1e3 rho_floor),
 * angular momentum (sum of S_3), rho min, p min, rho max, p max (finite values)
 * and the largest finite v^2 = gamma_ij v^i v^j; counts[2] = atmosphere cells,
 * invalid cells.  The sums are added in another order than the reference's
 * loop; minima, maxima and counts are exact. */
int grhd_global_diagnostics(grhd_ctx *ctx, double *out, int64_t *counts);
/* calc_ppi_modes (output.f90:995-1049): out[1 + 4 M], M = 4 = c0, then for
 * m = 1..M: Re and Im of sum_k m_k exp(-i m phi_k), amplitude |.|/c0 and phase
 * -atan2(Im, Re)/m, where m_k is the mass of phi plane k above the density cut
 * (planes reduced on the device, modes summed over k on the host in the
 * reference's order).  GRHD_ERR_UNSUPPORTED when nz < 2. */
int grhd_ppi_modes(grhd_ctx *ctx, double *out);
/* calc_m_dot (output.f90:951-993): -sum over (j, k) of D (alpha v^x - beta^x)
 * dy dz on the probe cell, the first centre at or outside the event horizon
 * (initialization.f90:117-190); *probe (may be NULL) = its index.
 * GRHD_ERR_UNSUPPORTED for Minkowski metrics, M < 0.1 or no such cell. */
int grhd_accretion_rate(grhd_ctx *ctx, double *mdot, int32_t *probe);

/* Unit entry points for parity checks. */
/* set_boundary_conditions on the device primitives (conditions.f90:773-1066). */
int grhd_boundaries(grhd_ctx *ctx);
/* calc_rhs (evolution.f90:30-191) of the current primitives (boundaries are
 * applied first); rhs has the interior conserved layout. */
int grhd_rhs(grhd_ctx *ctx, double *rhs);
/* cons_to_prim over all interior cells (grhd2.f90:252-260); repairs U in place. */
int grhd_con2prim(grhd_ctx *ctx);
/* Counters since creation: [0] cells reset to atmosphere by cons_to_prim,
 * [1] cells whose root solve fell back to Brent, [2] steps taken. */
int grhd_counters(grhd_ctx *ctx, int64_t *counters);

/* Stream and synchronisation: all work of a context is queued on its stream. */
int grhd_set_stream(grhd_ctx *ctx, void *hip_stream);
int grhd_synchronize(grhd_ctx *ctx);
/* Time of the last grhd_step / grhd_advance call measured with HIP events on
 * the context stream, in milliseconds, and the number of fused-stage kernel
 * launches it made. */
int grhd_last_timing(grhd_ctx *ctx, double *ms, int64_t *launches);

/* Stage form of a 3D problem above 65,536 cells (no counterpart in the
 * reference, whose strip loops have one structure, evolution.f90:48-189).
 * -1 (the default): chosen on the first step by timing one stage of each form;
 * 0: the split form (face fluxes of every direction into their arrays, then
 * the cell kernel); n > 0: the z-march form, in which each thread walks
 * n-cell segments of one column and differences the z fluxes in registers;
 * -2: the x-fused form, in which the cell kernel computes the x faces itself
 * (256 consecutive faces per block, differenced through LDS).  The z-march
 * and x-fused forms allocate a second primitive array.  Every form gives the
 * same results bit for bit.  GRHD_ERR_UNSUPPORTED for n > 0 or -2 on other
 * grids. */
int grhd_set_stage_form(grhd_ctx *ctx, int32_t zseg);

/* Face-kernel form of every direction of a problem outside the fused stage
 * form (no counterpart in the reference, whose face fluxes are computed in
 * its strip loops, evolution.f90:48-189).  -1 (the default): chosen on the
 * first step by timing each form on the problem; 1, 2, 4 or 8: that many
 * waves share an LDS copy of the face geometry; 0: the geometry read from a
 * per-problem table; 16: the table form with deferred faces -- for the MP5 +
 * sensor + HLLC scheme on the Kerr-Schild metrics, faces whose HLLC solve
 * falls back to HLLE are computed by a second kernel over a list of them;
 * 32 and 48: forms 0 and 16 with the x- and y-face blocks walking z first
 * (the z faces always do); 64 and 80: forms 0 and 16 with those blocks
 * walking bands of plane tiles (an eighth of the planes, one band per XCD)
 * per column tile.  Every form
 * gives the same results bit for bit.  GRHD_ERR_UNSUPPORTED for the table
 * forms (0, 16, 32, 48, 64, 80) where the problem has no tables, and for 16,
 * 48 or 80 where it has no deferred-face lists. */
int grhd_set_face_form(grhd_ctx *ctx, int32_t form);

/* Device bytes grhd_create allocates for cfg: the ghosted primitives, two
 * conserved buffers and the face fluxes (the reference allocates p, u, up and
 * rhs, initialization.f90:211-215, plus its metric cache).  Used to admit
 * problems into a batch by HBM footprint. */
int grhd_footprint(const grhd_config *cfg, uint64_t *bytes);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): grhd_step on a problem whose state the caller keeps in
 * its own device buffers, with no context.  nsteps SSP-RK3 steps at a fixed
 * dt (grhd2.f90:232-265: ghost fill, then three stages of calc_rhs, RK
 * combination, cons_to_prim and ghost fill), results bit-identical to
 * grhd_step.  Every intermediate (face fluxes, the RK stage buffer, the
 * coordinate table, the Michel injector, the counters) lives in one
 * file-scope device array of GRHD_NOALLOC_SCRATCH_BYTES per device, held until
 * the process ends; the call allocates nothing.
 *   p: device ghosted primitives, read and written, SoA [5][NZ][NY][NX] with
 *      NX = nx + 2 ng + 1 and NY (NZ) = ny + 2 ng + 1 (nz + 2 ng + 1) when
 *      ny (nz) > 1, else 1; ng = 1 Godunov, 2 TVD and WENO3, 3 MP5 and WENO5.
 *      For Michel (problem 7) the outer radial ghost layers must hold the
 *      injector state, as the initial data leaves them.
 *   u: device interior conserved variables U^n, read and written, [5][nz][ny][nx].
 *   stream: a hipStream_t (NULL: the default stream); the call returns after
 *      queueing the steps (the coordinate-table upload synchronises it).
 * GRHD_ERR_CAPACITY when the intermediates exceed the array (about 400k cells
 * in FP64 3D).  Calls that overlap on one device race on the array.  The
 * counters are cleared at each call and not reported. */
#define GRHD_NOALLOC_SCRATCH_BYTES (64u << 20)
int grhd_step_noalloc(const grhd_config *cfg, double *p, double *u, double dt, int32_t nsteps,
                      void *stream);

/* ------------------------------------------------------------------ batch --
 * Many independent problems resident on one device and advanced together.
 * The reference runs one problem per process (grhd2.f90); these entry points
 * have no counterpart there and serve the serving loop
 * (grhd_amd/serving.py): problems of different kinds and grid sizes are
 * admitted while others run, each keeps its own state and time, and every RK3
 * stage of all of them goes into shared launches.  A problem's results are
 * bit-identical to running it alone through grhd_advance. */
typedef struct grhd_batch grhd_batch;
/* capacity: slots; hbm_budget: bytes of problem state admitted at most
 * (0: 90 % of the free device memory). */
int grhd_batch_create(int32_t capacity, uint64_t hbm_budget, grhd_batch **out);
int grhd_batch_destroy(grhd_batch *b);
const char *grhd_batch_last_error(const grhd_batch *b);
/* Admits a problem: validation, device state, initial conditions (as
 * grhd_create + grhd_initial_conditions), its CFL dt.  GRHD_ERR_CAPACITY when
 * no slot is free or its footprint exceeds the remaining budget. */
int grhd_batch_add(grhd_batch *b, const grhd_config *cfg, int32_t *slot);
/* The problem's context, for grhd_set_primitives / grhd_perturb /
 * grhd_get_conserved / grhd_get_primitives between steps (not grhd_step). */
int grhd_batch_problem(grhd_batch *b, int32_t slot, grhd_ctx **ctx);
/* Frees a problem's slot and device state. */
int grhd_batch_remove(grhd_batch *b, int32_t slot);
/* Advances every unfinished problem by nsteps steps of the reference time
 * loop (grhd2.f90:179-265: dt, final-time clamp, three stages), each with
 * its own dt and time; a problem stops at its final_time (with pacing,
 * cheaper groups take more steps: grhd_batch_set_pacing). */
int grhd_batch_step(grhd_batch *b, int32_t nsteps);
/* done: final_time reached; t: its time; steps: steps taken in the batch. */
int grhd_batch_status(grhd_batch *b, int32_t slot, int32_t *done, double *t, int64_t *steps);
/* Launch form of the batched stage: 0 autotuned over forms 1-3 (default);
 * 1 one kernel per face direction, update, con2prim and ghost fill; 2-3 fused
 * faces, cells and ghost fill with 4/1 waves per face-geometry column; 4-5 one
 * persistent kernel per stage with 4/1 (selectable only: slower than the
 * phase forms).  All forms give identical results. */
int grhd_batch_set_variant(grhd_batch *b, int32_t variant);
/* Pacing (no reference counterpart; default 1).  With max_ratio > 1 and the
 * groups of different metric kinds or arithmetic on their own streams (the
 * phase launch forms), a group whose estimated step is at least 3 times
 * cheaper than the most expensive group's takes up to max_ratio (<= 8) steps
 * per step of that group in each chunk, so small problems are not held to the
 * step rate of the largest.
 * grhd_batch_step(nsteps) then advances the most expensive group nsteps steps
 * and cheaper ones up to max_ratio x nsteps; every problem still stops at its
 * final time, grhd_batch_status reports its steps, and its state equals its
 * state alone after that many steps, bit for bit. */
int grhd_batch_set_pacing(grhd_batch *b, int32_t max_ratio);
/* variant in use, unfinished problems, admitted bytes, device time (ms) and
 * kernel launches of the last grhd_batch_step. */
int grhd_batch_info(grhd_batch *b, int32_t *variant, int32_t *active, uint64_t *hbm_used,
                    double *last_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif
