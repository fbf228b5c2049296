/*
 * grhd_oracle.h -- CPU restatement of the GRHD2 per-step finite-volume update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle: only tests/, the
 * smoke entry point and the bench CPU-baseline leg may load it.  The product
 * path (libgrhd_hip.so) never links or calls it.
 *
 * Layout: the reference's own (Fortran) order.  Primitive array
 * p[((k*NYg + j)*NXg + i)*5 + v] with ghosted extents NXg = nx+2ng+1 (index
 * i_ref = i - ng runs over [-ng, nx+ng]); conserved arrays
 * u[(((k*ny + j)*nx + i)*5 + v] over interior cells only.  This is exactly the
 * memory image of p(neq,-ng:nx+ng,-ng:ny+ng,-ng:nz+ng) and u(neq,nx,ny,nz).
 */
#ifndef GRHD_ORACLE_H
#define GRHD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_MINKOWSKI = 0, ORC_EF = 1, ORC_KS = 2 };
enum { ORC_CARTESIAN = 0, ORC_CYLINDRICAL = 1, ORC_SPHERICAL = 2, ORC_SPHEROIDAL = 3 };
enum { ORC_REC_GODUNOV = 1, ORC_REC_TVD = 2, ORC_REC_WENO3 = 3, ORC_REC_MP5 = 4, ORC_REC_WENO5 = 5 };
enum { ORC_RS_HLLE = 1, ORC_RS_HLLC = 2 };
enum { ORC_LIM_MINMOD = 1, ORC_LIM_SUPERBEE = 2, ORC_LIM_MC = 3 };

/* Problem description: the .par surface the update path reads. */
typedef struct {
    int case_id;          /* parameters.f90:893-921 problem ids 1..12 */
    int metric, geom, log_r;
    double gam, mass, spin;
    int nx, ny, nz;
    double r_min, r_max, y_min, y_max, z_min, z_max;
    int rec, rs, lim, sensor;
    double cfl;
    /* boundary-data parameters used by set_boundary_conditions */
    double jet_ambient_density, jet_ambient_pressure, jet_ambient_velocity;
    double jet_nozzle_radius, jet_density, jet_pressure, jet_velocity;
    double dust_accretion_constant;
} orc_config;

typedef struct orc_problem orc_problem;

orc_problem *orc_create(const orc_config *cfg);
void orc_destroy(orc_problem *P);
/* dims[0..3] = nx, ny, nz, ng */
void orc_dims(const orc_problem *P, int *dims);
/* grid[0..8] = dx, dy, dz, x_min, x_max, y_min, y_max, z_min, z_max */
void orc_grid(const orc_problem *P, double *grid);

/* Michel outer injector (conditions.f90:229-246): ng x ny x nz x 5 values
 * copied from p(:, nx+1:nx+ng, 1:ny, 1:nz) of the given ghosted array. */
void orc_cache_injector(orc_problem *P, const double *p);

/* prim_to_cons over interior cells (grhd2.f90:126-134). */
void orc_prim_to_cons_all(const orc_problem *P, const double *p, double *u);
/* cons_to_prim over interior cells; u repaired in place (grhd2.f90:252-260). */
/* out[0] = calls of the Brent fallback of the con2prim root solve
 * (equations.f90:337-339, 343-424) and out[1] = iterations of its loop since
 * the last reset (process-wide; reset = 1 clears them after reading). */
void orc_brent_stats(unsigned long long *out, int reset);
/* out[0] = HLLC calls, out[1] = their HLLE fallbacks (fluxes.f90:174-397)
 * since the last reset (process-wide). */
void orc_hllc_stats(unsigned long long *out, int reset);
void orc_cons_to_prim_all(const orc_problem *P, double *u, double *p);
/* set_boundary_conditions (conditions.f90:773-1066). */
void orc_boundaries(const orc_problem *P, double *p);
/* calc_rhs (evolution.f90:30-191); rhs over interior cells. */
void orc_rhs(const orc_problem *P, const double *p, double *rhs);
/* update_adaptive_dt (evolution.f90:198-275); returns dt or -1 on invalid metric. */
double orc_dt(const orc_problem *P);
/* One SSP-RK3 step (grhd2.f90:229-265) with a given dt; u holds U^n on entry
 * and U^{n+1} on exit, p is overwritten with the new primitives. */
void orc_step(const orc_problem *P, double dt, double *u, double *p, double *scratch_u0,
              double *scratch_rhs);

/* Point probes. */
/* out[34], chris[64], dg[48]: same packing as the reference harness
 * (alpha, beta(3), gamma(3,3) column-major, det, gmunu(0:3,0:3), dlnalpha(0:3)). */
void orc_metric_point(const orc_problem *P, double x, double y, double *out, double *chris,
                      double *dg);
void orc_riemann(const orc_problem *P, const double *qL, const double *qR, int dir,
                 double alpha, const double *beta, const double *g, double sqg, double *flux);
void orc_physical_flux(const orc_problem *P, const double *q, int dir, double alpha,
                       const double *beta, const double *g, double sqg, double *flux);
/* prim[5*(nodes+2ng+1)] -> qL, qR[5*(nodes+1)] */
void orc_reconstruct(const orc_problem *P, const double *prim, int nodes, int dir, double *qL,
                     double *qR);
/* cons_to_prim / prim_to_cons / sources at interior cell (i,j,k), 1-based. */
void orc_cons_to_prim_cell(const orc_problem *P, double *cons, double *prim, int i, int j,
                           int k);
void orc_prim_to_cons_cell(const orc_problem *P, const double *prim, double *cons, int i, int j,
                           int k);
void orc_sources_cell(const orc_problem *P, const double *prim, double *src, int i, int j,
                      int k);

#ifdef __cplusplus
}
#endif
#endif
