// grhd_capi.hip -- the C ABI (include/grhd_c.h): contexts, validation, layout
// conversion between the reference's host layout and the device SoA layout,
// and the RK3 step driver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/grhd_c.h"
#include "grhd_ctx.hpp"
#include "grhd_initial.hpp"
#include "grhd_kernels.hpp"

using namespace grhd;


namespace {

std::mutex g_err_mu;
std::string g_create_err;

int fail(grhd_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) {
        ctx->err = msg;
    } else {
        std::lock_guard<std::mutex> lk(g_err_mu);
        g_create_err = msg;
    }
    return code;
}

#define HIPCHK(ctx, call)                                                                      \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(ctx, e_ == hipErrorOutOfMemory ? GRHD_ERR_OUT_OF_MEMORY : GRHD_ERR_HIP, \
                        std::string(#call) + ": " + hipGetErrorString(e_));                    \
    } while (0)

const double kPi = 3.141592653589793;

// validate_parameters (parameters.f90:359-485) for the fields the update reads.
int validate(const grhd_config &c, std::string &msg)
{
    const double eps = DBL_EPSILON;
    const bool polar = (c.geometry == GRHD_GEOM_SPHERICAL || c.geometry == GRHD_GEOM_SPHEROIDAL);
    const double reals[] = {c.r_min, c.r_max, c.y_min, c.y_max, c.z_min, c.z_max,
                            c.adiabatic_index, c.bh_mass, c.spin, c.cfl};
    for (double v : reals)
        if (!std::isfinite(v)) return msg = "real parameters must be finite", 1;
    if (c.problem < 1 || c.problem > 12) return msg = "invalid problem id", 1;
    if (c.nx < 1 || c.ny < 1 || c.nz < 1) return msg = "nx, ny and nz must be positive", 1;
    if (c.r_max <= c.r_min) return msg = "r_max must exceed r_min", 1;
    if (c.log_r && !polar) return msg = "logarithmic r needs spherical or spheroidal geometry", 1;
    if (polar && c.r_min <= 0.0) return msg = "spherical/spheroidal geometry needs r_min > 0", 1;
    if (c.adiabatic_index <= 1.0) return msg = "adiabatic_index must exceed 1", 1;
    if (c.bh_mass < 0.0) return msg = "bh_mass cannot be negative", 1;
    if (c.cfl <= 0.0) return msg = "cfl must be positive", 1;
    if (c.reconstruction < GRHD_REC_GODUNOV || c.reconstruction > GRHD_REC_WENO5)
        return msg = "invalid reconstruction identifier", 1;
    if (c.riemann_solver != GRHD_RS_HLLE && c.riemann_solver != GRHD_RS_HLLC)
        return msg = "invalid Riemann solver identifier", 1;
    if (c.reconstruction == GRHD_REC_TVD &&
        (c.tvd_limiter < GRHD_LIM_MINMOD || c.tvd_limiter > GRHD_LIM_MC))
        return msg = "invalid TVD limiter identifier", 1;
    if (c.precision != GRHD_PRECISION_FP64 && c.precision != GRHD_PRECISION_MIXED)
        return msg = "invalid precision", 1;
    switch (c.metric) {
    case GRHD_METRIC_MINKOWSKI:
        if (std::fabs(c.spin) > 0.0 || std::fabs(c.bh_mass) > 0.0)
            return msg = "Minkowski needs spin = 0 and bh_mass = 0", 1;
        if (c.geometry == GRHD_GEOM_SPHEROIDAL)
            return msg = "Minkowski needs cartesian, cylindrical or spherical geometry", 1;
        break;
    case GRHD_METRIC_EDDINGTON_FINKELSTEIN:
        if (std::fabs(c.spin) > 0.0) return msg = "Eddington-Finkelstein needs spin = 0", 1;
        if (c.bh_mass <= 0.0 || c.geometry != GRHD_GEOM_SPHERICAL)
            return msg = "Eddington-Finkelstein needs bh_mass > 0 and spherical geometry", 1;
        break;
    case GRHD_METRIC_KERR_SCHILD:
        if (c.bh_mass <= 0.0 || !(c.geometry == GRHD_GEOM_SPHEROIDAL || c.geometry == GRHD_GEOM_SPHERICAL))
            return msg = "Kerr-Schild needs bh_mass > 0 and spheroidal geometry", 1;
        if (std::fabs(c.spin) > c.bh_mass + 64.0 * eps * std::max(1.0, c.bh_mass))
            return msg = "Kerr-Schild needs |spin| <= bh_mass", 1;
        break;
    default:
        return msg = "invalid metric", 1;
    }
    if (polar) {
        if (c.y_min < 0.0 || c.y_max > kPi || c.y_max <= c.y_min)
            return msg = "need 0 <= theta_min < theta_max <= pi", 1;
        if (c.z_max <= c.z_min || c.z_max - c.z_min > 2.0 * kPi + 64.0 * eps * kPi)
            return msg = "the azimuthal interval must have length in (0, 2 pi]", 1;
    } else if (c.y_max <= c.y_min || c.z_max <= c.z_min) {
        return msg = "coordinate bounds must increase", 1;
    }
    if (c.problem == 12) {
        const double tol = 128.0 * eps * kPi;
        if (c.ny <= 1 || c.nz <= 1 || c.nz % 2 != 0)
            return msg = "fishbone_3d needs ny, nz > 1 and even nz", 1;
        if (std::fabs((c.z_max - c.z_min) - 2.0 * kPi) > tol)
            return msg = "fishbone_3d needs phi length 2 pi", 1;
        if (!c.allow_open_theta && (std::fabs(c.y_min) > tol || std::fabs(c.y_max - kPi) > tol))
            return msg = "fishbone_3d needs theta = [0, pi]", 1;
    }
    return 0;
}

int metric_kind(const grhd_config &c)
{
    if (c.metric == GRHD_METRIC_MINKOWSKI) {
        if (c.geometry == GRHD_GEOM_CARTESIAN) return MK_MINK_CART;
        if (c.geometry == GRHD_GEOM_CYLINDRICAL) return MK_MINK_CYL;
        return c.log_r ? MK_MINK_SPH_LOG : MK_MINK_SPH;
    }
    if (c.metric == GRHD_METRIC_EDDINGTON_FINKELSTEIN) return c.log_r ? MK_EF_LOG : MK_EF;
    return c.log_r ? MK_KS_LOG : MK_KS;
}

Params make_params(const grhd_config &c)
{
    Params P;
    std::memset(&P, 0, sizeof(P));
    P.problem = c.problem;
    P.mk = metric_kind(c);
    P.nx = c.nx;
    P.ny = c.ny;
    P.nz = c.nz;
    P.ng = (c.reconstruction == GRHD_REC_GODUNOV) ? 1
           : (c.reconstruction == GRHD_REC_TVD || c.reconstruction == GRHD_REC_WENO3) ? 2 : 3;
    P.rec = c.reconstruction;
    P.rs = c.riemann_solver;
    P.lim = c.tvd_limiter;
    P.sensor = c.shock_sensor ? 1 : 0;
    P.polar = (c.geometry == GRHD_GEOM_SPHERICAL || c.geometry == GRHD_GEOM_SPHEROIDAL);
    P.cyl = (c.geometry == GRHD_GEOM_CYLINDRICAL);
    P.face_gw[0] = P.face_gw[1] = P.face_gw[2] = 4;
    if (c.problem == 12) {
        const double tol = 128.0 * DBL_EPSILON * kPi;
        P.open_lo = c.y_min > tol;
        P.open_hi = c.y_max < kPi - tol;
    }
    P.gam = c.adiabatic_index;
    P.g1 = c.adiabatic_index / (c.adiabatic_index - 1.0);
    P.mass = c.bh_mass;
    P.spin = c.spin;
    P.rho_floor = 1.0e-10;           // initialization.f90:454-455
    P.p_floor = P.rho_floor * 1.0e-3;
    const double x_min = c.log_r ? std::log(c.r_min) : c.r_min;
    const double x_max = c.log_r ? std::log(c.r_max) : c.r_max;
    P.x_min = x_min;
    P.dx = (x_max - x_min) / (double)std::max(1, c.nx);
    P.y_min = c.y_min;
    P.dy = (c.y_max - c.y_min) / (double)std::max(1, c.ny);
    P.z_min = c.z_min;
    P.dz = (c.z_max - c.z_min) / (double)std::max(1, c.nz);
    P.inv_dx = 1.0 / P.dx;
    P.inv_dy = 1.0 / P.dy;
    P.inv_dz = 1.0 / P.dz;
    // ghosts only on active axes; an inactive axis stores its single cell
    P.ox = P.ng;
    P.NXs = c.nx + 2 * P.ng + 1;
    P.oy = (c.ny > 1) ? P.ng : -1;
    P.NYs = (c.ny > 1) ? c.ny + 2 * P.ng + 1 : 1;
    P.oz = (c.nz > 1) ? P.ng : -1;
    P.NZs = (c.nz > 1) ? c.nz + 2 * P.ng + 1 : 1;
    P.vstride = (int64_t)P.NXs * P.NYs * P.NZs;
    P.jet_ambient_density = c.jet_ambient_density;
    P.jet_ambient_pressure = c.jet_ambient_pressure;
    P.jet_ambient_velocity = c.jet_ambient_velocity;
    P.jet_nozzle_radius = c.jet_nozzle_radius;
    P.jet_density = c.jet_density;
    P.jet_pressure = c.jet_pressure;
    P.jet_velocity = c.jet_velocity;
    P.dust_accretion_constant = c.dust_accretion_constant;
    return P;
}

// Reference-layout index helpers (host arrays).
struct RefLayout {
    int nx, ny, nz, ng, NXg, NYg, NZg;
    explicit RefLayout(const Params &P)
        : nx(P.nx), ny(P.ny), nz(P.nz), ng(P.ng), NXg(P.nx + 2 * P.ng + 1), NYg(P.ny + 2 * P.ng + 1),
          NZg(P.nz + 2 * P.ng + 1) {}
    size_t p(int v, int i, int j, int k) const
    {
        return ((((size_t)(k + ng) * NYg + (size_t)(j + ng)) * NXg + (size_t)(i + ng)) * NEQ) + v;
    }
    size_t u(int v, int i, int j, int k) const
    {
        return ((((size_t)(k - 1) * ny + (size_t)(j - 1)) * nx + (size_t)(i - 1)) * NEQ) + v;
    }
    size_t psize() const { return (size_t)NXg * NYg * NZg * NEQ; }
    size_t usize() const { return (size_t)nx * ny * nz * NEQ; }
};

// Storage range of reference index along an axis: active axes keep [-ng, n+ng],
// inactive axes keep index 1 only.
inline void axis_range(int n, int ng, int &lo, int &hi)
{
    if (n > 1) {
        lo = -ng;
        hi = n + ng;
    } else {
        lo = 1;
        hi = 1;
    }
}

void host_to_soa_p(const Params &P, const double *href, std::vector<double> &soa)
{
    RefLayout L(P);
    soa.assign((size_t)P.vstride * NEQ, 0.0);
    int ilo, ihi, jlo, jhi, klo, khi;
    axis_range(P.nx, P.ng, ilo, ihi);
    if (P.nx == 1) { ilo = -P.ng; ihi = 1 + P.ng; }  // x keeps ghosts always
    axis_range(P.ny, P.ng, jlo, jhi);
    axis_range(P.nz, P.ng, klo, khi);
    for (int v = 0; v < NEQ; v++)
        for (int k = klo; k <= khi; k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = ilo; i <= ihi; i++)
                    soa[(size_t)v * P.vstride + (size_t)pindex(P, i, j, k)] = href[L.p(v, i, j, k)];
}

void soa_to_host_p(const Params &P, const std::vector<double> &soa, double *href)
{
    RefLayout L(P);
    const int ng = P.ng;
    for (int k = -ng; k <= P.nz + ng; k++) {
        const int ks = (P.nz > 1) ? k : 1;
        for (int j = -ng; j <= P.ny + ng; j++) {
            const int js = (P.ny > 1) ? j : 1;
            for (int i = -ng; i <= P.nx + ng; i++)
                for (int v = 0; v < NEQ; v++)
                    href[L.p(v, i, j, k)] = soa[(size_t)v * P.vstride + (size_t)pindex(P, i, js, ks)];
        }
    }
}

int ensure_dt(grhd_ctx *c)
{
    if (c->dt_ready) return GRHD_OK;
    const Params &P = c->P;
    if (!std::isfinite(P.dx) || P.dx <= 0.0 || (P.ny > 1 && (!std::isfinite(P.dy) || P.dy <= 0.0)) ||
        (P.nz > 1 && (!std::isfinite(P.dz) || P.dz <= 0.0)))
        return fail(c, GRHD_ERR_INVALID_METRIC, "invalid grid spacing in the CFL condition");
    const int nb = launch_cfl(P, c->stream, c->partial);
    std::vector<double> h((size_t)nb * 4);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->partial, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double mx = 0.0, my = 0.0, mz = 0.0, bad = 0.0;
    for (int b = 0; b < nb; b++) {
        mx = std::max(mx, h[4 * b + 0]);
        my = std::max(my, h[4 * b + 1]);
        mz = std::max(mz, h[4 * b + 2]);
        bad += h[4 * b + 3];
    }
    if (bad > 0.0) return fail(c, GRHD_ERR_INVALID_METRIC, "invalid metric data in CFL cells");
    double inv = mx / P.dx;
    if (P.ny > 1) inv = inv + my / P.dy;
    if (P.nz > 1) inv = inv + mz / P.dz;
    if (!std::isfinite(inv) || inv <= 0.0)
        return fail(c, GRHD_ERR_INVALID_METRIC, "invalid multidimensional CFL propagation rate");
    c->dt_cfl = c->cfg.cfl / inv;
    c->dt_ready = true;
    return GRHD_OK;
}

void cache_injector_from(grhd_ctx *c, const double *href)
{
    if (c->P.problem != 7) return;
    const Params &P = c->P;
    RefLayout L(P);
    std::vector<double> inj((size_t)NEQ * P.ng * P.ny * P.nz);
    for (int k = 1; k <= P.nz; k++)
        for (int j = 1; j <= P.ny; j++)
            for (int g = 1; g <= P.ng; g++)
                for (int v = 0; v < NEQ; v++)
                    inj[((((size_t)(k - 1) * P.ny + (j - 1)) * P.ng + (g - 1)) * NEQ) + v] =
                        href[L.p(v, P.nx + g, j, k)];
    (void)hipMemcpyAsync(c->injector, inj.data(), inj.size() * sizeof(double), hipMemcpyHostToDevice,
                         c->stream);
    (void)hipStreamSynchronize(c->stream);
}

int upload_p(grhd_ctx *c, const double *href)
{
    std::vector<double> soa;
    host_to_soa_p(c->P, href, soa);
    HIPCHK(c, hipMemcpyAsync(c->p, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    cache_injector_from(c, href);
    return GRHD_OK;
}

// The coordinate table of P in order (layout at Params::coord): put(v) for
// each entry.  Evaluated on the host with the reference's libm.
template <class Put>
void coord_entries(const Params &P, Put put)
{
    auto put_x = [&](double x) {
        const Coord q = coord_of(P.mk, x, 0.0);
        put(q.x);
        put(q.r);
        put(q.e2);
    };
    auto put_y = [&](double y) {
        put(std::sin(y));
        put(std::cos(y));
    };
    for (int i = -P.ng; i <= P.nx + P.ng; i++) put_x(xc(P, i));
    for (int i = 0; i <= P.nx; i++) put_x(xf(P, i));
    for (int j = -P.ng; j <= P.ny + P.ng; j++) put_y(yc(P, j));
    for (int j = 0; j <= P.ny; j++) put_y(yf(P, j));
}

// Face-kernel tiling of the problem, chosen once per context by timing: for
// each active direction, the face kernel alone with 4, 2, 8 and 1 waves sharing
// one LDS copy of the face geometry, with the geometry read from a
// per-problem table, and with the table and the HLLC fallbacks deferred to a
// second kernel (one warm-up launch, then three timed; two rounds), on the
// problem's own primitives.  The face kernels write only the face arrays (and
// the table), which the next stage rewrites, and every form gives the same
// fluxes, so tuning changes no result.  Grids in the fused stage form keep
// groups of four.
// GRHD_TRACE_TUNE=1 prints the tuner's timings to stderr.
bool trace_tune()
{
    static const bool on = [] {
        const char *e = std::getenv("GRHD_TRACE_TUNE");
        return e && e[0] == '1';
    }();
    return on;
}

int tune_stage_form(grhd_ctx *c, double dt);

int tune_faces(grhd_ctx *c, double dt)
{
    c->faces_tuned = true;
    if (single_stage_fused(c->P)) return GRHD_OK;
    static const int cand[] = {4, 2, 8, 1, 0, kFaceTableDefer, kFaceOrderZ, kFaceTableDefer | kFaceOrderZ,
                               kFaceOrderBand, kFaceTableDefer | kFaceOrderBand};
    for (int d = 0; d < 3; d++) {
        if (!(d == 0 || (d == 1 && c->P.ny > 1) || (d == 2 && c->P.nz > 1))) continue;
        float best = 0.0f;
        int best_gw = 4;
        // two rounds over the candidates, the fastest time of each kept: the
        // clock rises over the first few hundred milliseconds of load, which
        // would favour the candidates timed last, and the first round of a
        // table form also fills its table
        for (int round = 0; round < 2; round++)
        for (int gw : cand) {
            const bool table = face_form_table(gw);
            if (table && !c->geo[d]) continue;
            if (face_form_defers(gw) && !c->defer_list[d]) continue;
            // the z faces walk z first in every order
            if (d == 2 && (gw & kFaceOrders)) continue;
            // a forced form: only that candidate (its table filled by the
            // warm-up launch; the z faces take a z-first form without the flag)
            if (c->forced_face >= 0 && gw != (d == 2 ? c->forced_face & ~kFaceOrders : c->forced_face)) continue;
            Params P = c->P;
            P.face_gw[d] = gw;
            P.geo_tab[d] = table ? c->geo[d] : nullptr;
            P.geo_ready = c->geo_filled[d] ? 1 : 0;  // the warm-up launch fills the table
            if (table) c->geo_filled[d] = true;
            for (int rep = 0; rep < 4; rep++) {
                if (rep == 1) HIPCHK(c, hipEventRecord(c->ev0, c->stream));
                if (c->mixed) launch_face_dir_f32(P, c->stream, c->p, c->F, c->fstr, d + 1);
                else launch_face_dir_f64(P, c->stream, c->p, c->F, c->fstr, d + 1);
            }
            HIPCHK(c, hipEventRecord(c->ev1, c->stream));
            HIPCHK(c, hipEventSynchronize(c->ev1));
            float ms = 0.0f;
            HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
            if (trace_tune()) {
                std::fprintf(stderr, "grhd tune: direction %d, %s %d: %.4f ms per launch\n", d + 1,
                             !face_form_table(gw) ? "waves per geometry group"
                             : face_form_defers(gw) ? "geometry table, deferred faces (+32 z first, +64 bands)"
                                                    : "geometry table (+32 z first, +64 bands)",
                             gw, ms / 3.0);
                std::fflush(stderr);
            }
            if (best == 0.0f || ms < best) {
                best = ms;
                best_gw = gw;
            }
        }
        c->P.face_gw[d] = best_gw;
        c->P.geo_tab[d] = face_form_table(best_gw) ? c->geo[d] : nullptr;
    }
    HIPCHK(c, hipGetLastError());
    const int rc = tune_stage_form(c, dt);
    if (rc != GRHD_OK) return rc;
    // the stage skips the geometry kernels once every active direction's
    // table is filled (a forced form other than the table forms leaves them
    // unfilled, and the z-march form then fills the z table every stage)
    bool filled = true;
    for (int d = 0; d < 3; d++)
        if ((d == 0 || (d == 1 && c->P.ny > 1) || (d == 2 && c->P.nz > 1)) && !c->geo_filled[d]) filled = false;
    c->P.geo_ready = filled ? 1 : 0;
    return GRHD_OK;
}

// Stage form of a 3D problem, chosen once per context after the face tiling
// by timing the three stages of one step (the step's dt) of each candidate:
// the split form (faces of every direction into their arrays, then k_cell)
// and the z-march form with 16-cell segments (8-cell segments measured the
// same on the 3D torus, 32-cell ones 25 % slower: too few waves).  The
// x-fused form (x faces computed by the cell kernel, k_xcell) is selected
// only by grhd_set_stage_form: over whole steps of the 3D torus it measured
// 0.5 % slower than the split form in FP64 and 2 % in mixed precision, while
// these stage timings, on scratch buffers, put it 4-6 % ahead.  Every form
// gives the same results bit for bit.  The timing runs write only the stage buffer u1, the
// face arrays and tables, the second primitive array and the tuner's
// counters: the split form runs on a copy of the primitives in p_alt.
bool zmarch_possible(const Params &P) { return !single_stage_fused(P) && P.ny > 1 && P.nz > 1; }

int tune_stage_form(grhd_ctx *c, double dt)
{
    const Params &P0 = c->P;
    c->P.zseg = 0;
    if (!zmarch_possible(P0) || !c->geo[2] || (c->forced_zseg == -1 && P0.nz < 16)) return GRHD_OK;
    const size_t pbytes = (size_t)P0.vstride * NEQ * sizeof(double);
    if (c->forced_zseg == 0) {
        c->P.zseg = 0;
        return GRHD_OK;
    }
    if (!c->p_alt) {
        const hipError_t e = hipMalloc(&c->p_alt, pbytes);
        if (e != hipSuccess) {  // no room for a second primitive array: split form
            (void)hipGetLastError();
            c->p_alt = nullptr;
            if (c->forced_zseg != -1) return fail(c, GRHD_ERR_OUT_OF_MEMORY, "second primitive array");
            return GRHD_OK;
        }
    }
    if (c->forced_zseg != -1) {  // a forced ping-pong form
        c->P.zseg = c->forced_zseg;
        return GRHD_OK;
    }
    double *u0 = c->u[c->ucur], *u1 = c->u[1 - c->ucur];
    unsigned long long *cnt = c->counters + 4;
    static const int cand[] = {0, 16};
    float best = 0.0f;
    int best_seg = 0;
    for (int round = 0; round < 2; round++)  // as in tune_faces
    for (int seg : cand) {
        if (seg > P0.nz) continue;
        void *geo = c->geo[2];
        Params P = P0;
        P.zseg = seg;
        if (seg == 0) HIPCHK(c, hipMemcpyAsync(c->p_alt, c->p, pbytes, hipMemcpyDeviceToDevice, c->stream));
        // a warm-up stage 1, then stages 1, 2 and 3 (stages 2 and 3 also
        // read the stage buffer): the mix of one step
        for (int rep = 0; rep < 4; rep++) {
            const int stage = (rep == 0) ? 1 : rep;
            if (rep == 1) HIPCHK(c, hipEventRecord(c->ev0, c->stream));
            if (seg == 0) launch_stage(P, c->stream, c->p_alt, c->F, c->fstr, u0, u1, stage, dt, cnt, c->mixed,
                                       c->injector, false);
            else launch_stage_z(P, c->stream, c->p, c->p_alt, c->F, c->fstr, geo, u0, u1, stage, dt, cnt,
                                c->mixed);
        }
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        HIPCHK(c, hipEventSynchronize(c->ev1));
        float ms = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        if (trace_tune()) {
            std::fprintf(stderr, "grhd tune: stage form %s %d: %.4f ms per stage\n",
                         seg == kStageXFused ? "x-fused" : seg ? "z-march segment" : "split", seg, ms / 3.0);
            std::fflush(stderr);
        }
        // a ping-pong form (a second primitive array and a copy per call)
        // only for a clear gain: 2 % faster than the best so far
        if (best == 0.0f || ms < (seg == 0 ? best : 0.98f * best)) {
            best = ms;
            best_seg = seg;
        }
    }
    HIPCHK(c, hipGetLastError());
    c->P.zseg = best_seg;
    if (best_seg == 0) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        (void)hipFree(c->p_alt);
        c->p_alt = nullptr;
    }
    return GRHD_OK;
}

int run_steps(grhd_ctx *c, const double *dts, int n)
{
    if (!c->faces_tuned && n > 0) {
        launch_boundaries(c->P, c->stream, c->p, c->injector);
        const int rc = tune_faces(c, dts[0]);
        if (rc != GRHD_OK) return rc;
    }
    const Params &P = c->P;
    // the ping-pong stage forms (z-march, x-fused) read one primitive array
    // and write the other
    const bool zmarch = P.zseg != 0 && c->p_alt;
    const bool xfused = zmarch && P.zseg == kStageXFused;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    int64_t launches = 0;
    launch_boundaries(P, c->stream, c->p, c->injector);  // grhd2.f90:233
    launches++;
    // the ping-pong forms write the stage's primitives to the other array; its
    // cells that no ghost fill rewrites (the outermost ghost layer) must hold
    // the same values in both
    if (zmarch && n > 0)
        HIPCHK(c, hipMemcpyAsync(c->p_alt, c->p, (size_t)P.vstride * NEQ * sizeof(double), hipMemcpyDeviceToDevice,
                                 c->stream));
    // launches per stage: the face kernels of each active direction (and the
    // geometry-table kernel of a direction in the table forms, until the
    // tables are filled, and the deferred-face kernel), the cell kernel (with
    // the z faces in the z-march form, the x faces in the x-fused form) and
    // the ghost fill
    int per_stage = 3;  // fused form: all faces, cells, ghosts
    if (!single_stage_fused(P)) {
        per_stage = 2;
        for (int d = 0; d < 3; d++) {
            const bool on = (d == 0) || (d == 1 && P.ny > 1) || (d == 2 && P.nz > 1);
            if (!on || (zmarch && d == (xfused ? 0 : 2))) continue;
            const bool table = face_form_table(P.face_gw[d]);
            per_stage += 1 + (table && !P.geo_ready) + (face_form_defers(P.face_gw[d]) && P.defer_list[d]);
        }
        if (zmarch && !P.geo_ready) per_stage += 1;
    }
    for (int s = 0; s < n; s++) {
        double *u0 = c->u[c->ucur], *u1 = c->u[1 - c->ucur];
        for (int rk = 1; rk <= 3; rk++) {
            if (zmarch) {
                launch_stage_z(P, c->stream, c->p, c->p_alt, c->F, c->fstr, c->geo[xfused ? 0 : 2], u0, u1, rk,
                               dts[s], c->counters, c->mixed);
                launch_boundaries(P, c->stream, c->p_alt, c->injector);
                std::swap(c->p, c->p_alt);
            } else {
                launch_stage(P, c->stream, c->p, c->F, c->fstr, u0, u1, rk, dts[s], c->counters, c->mixed,
                             c->injector, true);
            }
            launches += per_stage;
        }
        c->ucur = 1 - c->ucur;
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->steps += n;
    c->last_launches = launches;
    return GRHD_OK;
}

}  // namespace

namespace grhd {
uint64_t ctx_footprint(const grhd_config &cfg)
{
    grhd_config c = cfg;
    if (c.metric == GRHD_METRIC_KERR_SCHILD) c.geometry = GRHD_GEOM_SPHEROIDAL;
    const Params P = make_params(c);
    const uint64_t ncell = (uint64_t)P.nx * P.ny * P.nz;
    const uint64_t rsz = (cfg.precision == GRHD_PRECISION_MIXED) ? sizeof(float) : sizeof(double);
    const uint64_t nf[3] = {(uint64_t)(P.nx + 1) * P.ny * P.nz,
                            P.ny > 1 ? (uint64_t)P.nx * (P.ny + 1) * P.nz : 1,
                            P.nz > 1 ? (uint64_t)P.nx * P.ny * (P.nz + 1) : 1};
    uint64_t b = (uint64_t)P.vstride * NEQ * sizeof(double) + 2 * ncell * NEQ * sizeof(double);
    for (int d = 0; d < 3; d++) b += nf[d] * (d == 0 || nf[d] > 1 ? NEQ : 1) * rsz;
    if (!single_stage_fused(P))  // face-geometry tables
        for (int d = 0; d < 3; d++)
            if (d == 0 || (d == 1 && P.ny > 1) || (d == 2 && P.nz > 1))
                b += (uint64_t)NFG * (P.nx + (d == 0)) * (P.ny + (d == 1)) * rsz;
    if (defer_faces(P))  // deferred-face lists and counters
        for (int d = 0; d < 3; d++) b += (nf[d] > 1 ? nf[d] : 0) * sizeof(unsigned) + 2 * sizeof(unsigned);
    b += (uint64_t)NEQ * P.ng * P.ny * P.nz * sizeof(double) + 8;
    b += (uint64_t)std::max<int64_t>((((int64_t)P.nx * P.ny + 255) / 256) * 4, std::max<int64_t>(4096, P.nz)) *
             sizeof(double) +
         8 * sizeof(unsigned long long);
    b += (uint64_t)coord_table_size(P.nx, P.ny, P.ng) * sizeof(double);
    return b;
}
}  // namespace grhd

extern "C" {

int grhd_footprint(const grhd_config *cfg, uint64_t *bytes)
{
    if (!cfg || !bytes) return fail(nullptr, GRHD_ERR_INVALID_ARGUMENT, "null argument");
    std::string msg;
    if (validate(*cfg, msg)) return fail(nullptr, GRHD_ERR_INVALID_CONFIG, msg);
    *bytes = ctx_footprint(*cfg);
    return GRHD_OK;
}

const char *grhd_last_error(const grhd_ctx *ctx)
{
    if (ctx) return ctx->err.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    return g_create_err.c_str();
}

int grhd_validate(const grhd_config *cfg)
{
    if (!cfg) return fail(nullptr, GRHD_ERR_INVALID_ARGUMENT, "null argument");
    std::string msg;
    if (validate(*cfg, msg)) return fail(nullptr, GRHD_ERR_INVALID_CONFIG, msg);
    const Params P = make_params(*cfg);
    if ((P.problem == 5 && (P.nx < P.ng || (P.nz > 1 && P.nz < P.ng))) ||
        (P.problem == 11 && P.ny < P.ng) || (P.problem == 12 && P.ny < P.ng))
        return fail(nullptr, GRHD_ERR_INVALID_CONFIG, "grid smaller than the ghost depth");
    return GRHD_OK;
}

int grhd_create(const grhd_config *cfg, grhd_ctx **out)
{
    if (!cfg || !out) return fail(nullptr, GRHD_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    std::string msg;
    if (validate(*cfg, msg)) return fail(nullptr, GRHD_ERR_INVALID_CONFIG, msg);
    grhd_ctx *c = new grhd_ctx();
    c->cfg = *cfg;
    if (c->cfg.metric == GRHD_METRIC_KERR_SCHILD) c->cfg.geometry = GRHD_GEOM_SPHEROIDAL;
    c->P = make_params(c->cfg);
    c->mixed = (cfg->precision == GRHD_PRECISION_MIXED);
    const Params &P = c->P;
    // Boundary rules that reflect or shift need as many interior cells as ghosts.
    const bool needs = (P.problem == 5 && (P.nx < P.ng || (P.nz > 1 && P.nz < P.ng))) || (P.problem == 11 && P.ny < P.ng) ||
                       (P.problem == 12 && P.ny < P.ng);
    if (needs) {
        delete c;
        return fail(nullptr, GRHD_ERR_INVALID_CONFIG, "grid smaller than the ghost depth");
    }
    auto bail = [&](hipError_t e, const char *what) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        const int code = (e == hipErrorOutOfMemory) ? GRHD_ERR_OUT_OF_MEMORY : GRHD_ERR_HIP;
        grhd_destroy(c);
        return fail(nullptr, code, m);
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(e, "hipStreamCreate");
    c->own_stream = true;
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    const size_t rsz = c->mixed ? sizeof(float) : sizeof(double);
    c->fstr[0] = (int64_t)(P.nx + 1) * P.ny * P.nz;
    c->fstr[1] = (int64_t)P.nx * (P.ny + 1) * P.nz;
    c->fstr[2] = (int64_t)P.nx * P.ny * (P.nz + 1);
    if ((e = hipMalloc(&c->p, (size_t)P.vstride * NEQ * sizeof(double))) != hipSuccess) return bail(e, "hipMalloc p");
    for (int b = 0; b < 2; b++)
        if ((e = hipMalloc(&c->u[b], (size_t)ncell * NEQ * sizeof(double))) != hipSuccess)
            return bail(e, "hipMalloc u");
    for (int d = 0; d < 3; d++) {
        const bool active = (d == 0) || (d == 1 && P.ny > 1) || (d == 2 && P.nz > 1);
        const size_t n = active ? (size_t)c->fstr[d] * NEQ : 1;
        if ((e = hipMalloc(&c->F[d], n * rsz)) != hipSuccess) return bail(e, "hipMalloc flux");
    }
    if (!single_stage_fused(P)) {
        for (int d = 0; d < 3; d++) {
            const bool active = (d == 0) || (d == 1 && P.ny > 1) || (d == 2 && P.nz > 1);
            if (!active) continue;
            const size_t nab = (size_t)(P.nx + (d == 0)) * (P.ny + (d == 1));
            if ((e = hipMalloc(&c->geo[d], (size_t)NFG * nab * rsz)) != hipSuccess) return bail(e, "hipMalloc geometry");
        }
    }
    if (defer_faces(P)) {
        for (int d = 0; d < 3; d++) {
            const bool active = (d == 0) || (d == 1 && P.ny > 1) || (d == 2 && P.nz > 1);
            if (!active) continue;
            if ((e = hipMalloc(&c->defer_list[d], (size_t)c->fstr[d] * sizeof(unsigned))) != hipSuccess)
                return bail(e, "hipMalloc deferred faces");
            c->P.defer_list[d] = c->defer_list[d];
        }
        if ((e = hipMalloc(&c->defer_ctr, 6 * sizeof(unsigned))) != hipSuccess)
            return bail(e, "hipMalloc deferred-face counters");
        if ((e = hipMemsetAsync(c->defer_ctr, 0, 6 * sizeof(unsigned), c->stream)) != hipSuccess)
            return bail(e, "hipMemset");
        c->P.defer_ctr = c->defer_ctr;
    }
    if ((e = hipMalloc(&c->injector, (size_t)NEQ * P.ng * P.ny * P.nz * sizeof(double) + 8)) != hipSuccess)
        return bail(e, "hipMalloc injector");
    c->partial_n = (int)(((int64_t)P.nx * P.ny + 255) / 256);
    c->partial_cap = std::max<int64_t>((int64_t)c->partial_n * 4, std::max<int64_t>(4096, P.nz));
    if ((e = hipMalloc(&c->partial, (size_t)c->partial_cap * sizeof(double))) != hipSuccess)
        return bail(e, "hipMalloc partial");
    if ((e = hipMalloc(&c->counters, 8 * sizeof(unsigned long long))) != hipSuccess)
        return bail(e, "hipMalloc counters");
    {
        // coordinate table: the grid's x, r, e^{2x}, sin and cos evaluated once
        // on the host (initialization.f90:84-263)
        std::vector<double> t((size_t)coord_table_size(P.nx, P.ny, P.ng));
        size_t o = 0;
        coord_entries(P, [&](double v) { t[o++] = v; });
        if ((e = hipMalloc(&c->coord, t.size() * sizeof(double))) != hipSuccess) return bail(e, "hipMalloc coord");
        if ((e = hipMemcpy(c->coord, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
            return bail(e, "hipMemcpy coord");
        c->P.coord = c->coord;
    }
    if ((e = hipMemsetAsync(c->counters, 0, 8 * sizeof(unsigned long long), c->stream)) != hipSuccess)
        return bail(e, "hipMemset");
    if ((e = hipMemsetAsync(c->p, 0, (size_t)P.vstride * NEQ * sizeof(double), c->stream)) != hipSuccess)
        return bail(e, "hipMemset");
    if ((e = hipEventCreate(&c->ev0)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreate(&c->ev1)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail(e, "hipStreamSynchronize");
    *out = c;
    return GRHD_OK;
}

int grhd_destroy(grhd_ctx *c)
{
    if (!c) return GRHD_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->p);
    (void)hipFree(c->p_alt);
    (void)hipFree(c->u[0]);
    (void)hipFree(c->u[1]);
    for (int d = 0; d < 3; d++) {
        (void)hipFree(c->F[d]);
        (void)hipFree(c->geo[d]);
        (void)hipFree(c->defer_list[d]);
    }
    (void)hipFree(c->defer_ctr);
    (void)hipFree(c->injector);
    (void)hipFree(c->partial);
    (void)hipFree(c->counters);
    (void)hipFree(c->coord);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GRHD_OK;
}

int grhd_dims(const grhd_ctx *c, int32_t *dims)
{
    if (!c || !dims) return GRHD_ERR_INVALID_ARGUMENT;
    dims[0] = c->P.nx;
    dims[1] = c->P.ny;
    dims[2] = c->P.nz;
    dims[3] = c->P.ng;
    return GRHD_OK;
}

int grhd_initial_conditions(grhd_ctx *c)
{
    if (!c) return GRHD_ERR_INVALID_ARGUMENT;
    std::vector<double> h;
    const int rc = initial_conditions(c->P, c->cfg, h, c->err);
    if (rc != GRHD_OK) return rc;
    return grhd_set_primitives(c, h.data());
}

int grhd_perturb(grhd_ctx *c, int32_t type, double amplitude, uint64_t seed)
{
    if (!c || (type < 1 || type > 3) || !(amplitude >= 0.0 && amplitude < 1.0))
        return c ? fail(c, GRHD_ERR_INVALID_ARGUMENT, "perturbation type must be 1, 2 or 3 and 0 <= A < 1")
                 : GRHD_ERR_INVALID_ARGUMENT;
    RefLayout L(c->P);
    std::vector<double> h(L.psize());
    int rc = grhd_get_primitives(c, h.data());
    if (rc != GRHD_OK) return rc;
    perturb(c->P, type, amplitude, seed, h);
    rc = upload_p(c, h.data());
    if (rc != GRHD_OK) return rc;
    launch_p2c(c->P, c->stream, c->p, c->u[c->ucur]);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GRHD_OK;
}

int grhd_fishbone_unit_density_k(const grhd_config *cfg, double *K)
{
    if (!cfg || !K) return GRHD_ERR_INVALID_ARGUMENT;
    grhd_config c = *cfg;
    if (c.metric == GRHD_METRIC_KERR_SCHILD) c.geometry = GRHD_GEOM_SPHEROIDAL;
    const Params P = make_params(c);
    const double k = fishbone_K_for_unit_density(P, c.fm_inner_radius * c.bh_mass,
                                                 c.fm_pressure_max_radius * c.bh_mass);
    if (!(k > 0.0) || !std::isfinite(k)) return fail(nullptr, GRHD_ERR_INVALID_CONFIG, "invalid torus");
    *K = k;
    return GRHD_OK;
}

extern "C++" {
namespace {
// alpha and beta^i at the interior cell centres, from the same metric
// functions and coordinate evaluation the coordinate table uses
// (precalculate_metric_cache, initialization.f90:283-299).  The metrics are
// independent of the third coordinate, so one (i, j) column fills every k.
template <int MK>
void lapse_shift_mk(const Params &P, double *alpha, double *beta)
{
    for (int j = 1; j <= P.ny; j++) {
        const double y = yc(P, j);
        for (int i = 1; i <= P.nx; i++) {
            FaceMetric m;
            metric_face<MK>(P, coord_of(P.mk, xc(P, i), y), m);
            for (int k = 0; k < P.nz; k++) {
                const int64_t t = ((int64_t)k * P.ny + (j - 1)) * P.nx + (i - 1);
                if (alpha) alpha[t] = m.alpha;
                if (beta)
                    for (int d = 0; d < 3; d++) beta[3 * t + d] = m.beta[d];
            }
        }
    }
}
}  // namespace
}  // extern "C++"

int grhd_lapse_shift(const grhd_config *cfg, double *alpha, double *beta)
{
    if (!cfg || (!alpha && !beta)) return fail(nullptr, GRHD_ERR_INVALID_ARGUMENT, "null argument");
    std::string msg;
    if (validate(*cfg, msg)) return fail(nullptr, GRHD_ERR_INVALID_CONFIG, msg);
    grhd_config c = *cfg;
    if (c.metric == GRHD_METRIC_KERR_SCHILD) c.geometry = GRHD_GEOM_SPHEROIDAL;
    const Params P = make_params(c);
    switch (P.mk) {
    case MK_MINK_CART: lapse_shift_mk<MK_MINK_CART>(P, alpha, beta); break;
    case MK_MINK_CYL: lapse_shift_mk<MK_MINK_CYL>(P, alpha, beta); break;
    case MK_MINK_SPH: lapse_shift_mk<MK_MINK_SPH>(P, alpha, beta); break;
    case MK_MINK_SPH_LOG: lapse_shift_mk<MK_MINK_SPH_LOG>(P, alpha, beta); break;
    case MK_EF: lapse_shift_mk<MK_EF>(P, alpha, beta); break;
    case MK_EF_LOG: lapse_shift_mk<MK_EF_LOG>(P, alpha, beta); break;
    case MK_KS: lapse_shift_mk<MK_KS>(P, alpha, beta); break;
    default: lapse_shift_mk<MK_KS_LOG>(P, alpha, beta); break;
    }
    return GRHD_OK;
}

int grhd_put_primitives(grhd_ctx *c, const double *p)
{
    if (!c || !p) return GRHD_ERR_INVALID_ARGUMENT;
    return upload_p(c, p);
}

int grhd_set_primitives(grhd_ctx *c, const double *p)
{
    if (!c || !p) return GRHD_ERR_INVALID_ARGUMENT;
    int rc = upload_p(c, p);
    if (rc != GRHD_OK) return rc;
    launch_p2c(c->P, c->stream, c->p, c->u[c->ucur]);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GRHD_OK;
}

int grhd_put_conserved(grhd_ctx *c, const double *u)
{
    if (!c || !u) return GRHD_ERR_INVALID_ARGUMENT;
    const Params &P = c->P;
    RefLayout L(P);
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    std::vector<double> soa((size_t)ncell * NEQ);
    for (int64_t t = 0; t < ncell; t++)
        for (int v = 0; v < NEQ; v++) soa[(size_t)v * ncell + t] = u[(size_t)t * NEQ + v];
    HIPCHK(c, hipMemcpyAsync(c->u[c->ucur], soa.data(), soa.size() * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GRHD_OK;
}

int grhd_set_conserved(grhd_ctx *c, const double *u)
{
    int rc = grhd_put_conserved(c, u);
    if (rc != GRHD_OK) return rc;
    if (c->P.problem == 7) {  // prepare_case_boundary_data (conditions.f90:217-226)
        std::vector<double> h;
        rc = initial_conditions(c->P, c->cfg, h, c->err);
        if (rc != GRHD_OK) return rc;
        cache_injector_from(c, h.data());
    }
    rc = grhd_con2prim(c);
    if (rc != GRHD_OK) return rc;
    return grhd_boundaries(c);
}

int grhd_get_primitives(grhd_ctx *c, double *p)
{
    if (!c || !p) return GRHD_ERR_INVALID_ARGUMENT;
    std::vector<double> soa((size_t)c->P.vstride * NEQ);
    HIPCHK(c, hipMemcpyAsync(soa.data(), c->p, soa.size() * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    soa_to_host_p(c->P, soa, p);
    return GRHD_OK;
}

int grhd_get_conserved(grhd_ctx *c, double *u)
{
    if (!c || !u) return GRHD_ERR_INVALID_ARGUMENT;
    const Params &P = c->P;
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    std::vector<double> soa((size_t)ncell * NEQ);
    HIPCHK(c, hipMemcpyAsync(soa.data(), c->u[c->ucur], soa.size() * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int64_t t = 0; t < ncell; t++)
        for (int v = 0; v < NEQ; v++) u[(size_t)t * NEQ + v] = soa[(size_t)v * ncell + t];
    return GRHD_OK;
}

int grhd_compute_dt(grhd_ctx *c, double *dt)
{
    if (!c || !dt) return GRHD_ERR_INVALID_ARGUMENT;
    const int rc = ensure_dt(c);
    if (rc != GRHD_OK) return rc;
    *dt = c->dt_cfl;
    return GRHD_OK;
}

int grhd_step(grhd_ctx *c, double dt, int32_t nsteps)
{
    if (!c || nsteps < 0 || !std::isfinite(dt)) return GRHD_ERR_INVALID_ARGUMENT;
    std::vector<double> dts((size_t)nsteps, dt);
    return run_steps(c, dts.data(), nsteps);
}

int grhd_advance(grhd_ctx *c, double t_final, int32_t max_steps, double *t, int32_t *steps_done,
                 double *last_dt)
{
    if (!c || !t || max_steps < 0) return GRHD_ERR_INVALID_ARGUMENT;
    int rc = ensure_dt(c);
    if (rc != GRHD_OK) return rc;
    // grhd2.f90:179-184,229-230: dt, clamp to the final time, t advances first.
    std::vector<double> dts;
    double tt = *t, dt = 0.0;
    while ((int)dts.size() < max_steps && tt < t_final) {
        dt = c->dt_cfl;
        if (tt + dt > t_final) dt = t_final - tt;
        tt = tt + dt;
        dts.push_back(dt);
    }
    rc = run_steps(c, dts.data(), (int)dts.size());
    if (rc != GRHD_OK) return rc;
    *t = tt;
    if (steps_done) *steps_done = (int32_t)dts.size();
    if (last_dt) *last_dt = dt;
    return GRHD_OK;
}

int grhd_boundaries(grhd_ctx *c)
{
    if (!c) return GRHD_ERR_INVALID_ARGUMENT;
    launch_boundaries(c->P, c->stream, c->p, c->injector);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GRHD_OK;
}

int grhd_rhs(grhd_ctx *c, double *rhs)
{
    if (!c || !rhs) return GRHD_ERR_INVALID_ARGUMENT;
    const Params &P = c->P;
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    double *tmp = nullptr;
    HIPCHK(c, hipMalloc(&tmp, (size_t)ncell * NEQ * sizeof(double)));
    launch_boundaries(P, c->stream, c->p, c->injector);
    launch_stage(P, c->stream, c->p, c->F, c->fstr, c->u[c->ucur], tmp, 0, 0.0, c->counters, c->mixed,
                 c->injector, false);
    std::vector<double> soa((size_t)ncell * NEQ);
    hipError_t e = hipMemcpyAsync(soa.data(), tmp, soa.size() * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(c, GRHD_ERR_HIP, hipGetErrorString(e));
    for (int64_t t = 0; t < ncell; t++)
        for (int v = 0; v < NEQ; v++) rhs[(size_t)t * NEQ + v] = soa[(size_t)v * ncell + t];
    return GRHD_OK;
}

int grhd_con2prim(grhd_ctx *c)
{
    if (!c) return GRHD_ERR_INVALID_ARGUMENT;
    launch_c2p(c->P, c->stream, c->p, c->u[c->ucur], c->counters);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GRHD_OK;
}

int grhd_total_mass(grhd_ctx *c, double *mass)
{
    if (!c || !mass) return GRHD_ERR_INVALID_ARGUMENT;
    const int nb = launch_mass(c->P, c->stream, c->u[c->ucur], c->partial, (int)std::min<int64_t>(c->partial_cap, 4096));
    HIPCHK(c, hipGetLastError());
    std::vector<double> h((size_t)nb);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->partial, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double m = 0.0;
    for (int b = 0; b < nb; b++) m = m + h[(size_t)b];
    *mass = m;
    return GRHD_OK;
}

int grhd_convergence_norms(grhd_ctx *c, double t, double *norms, int32_t *imax)
{
    if (!c || !norms) return GRHD_ERR_INVALID_ARGUMENT;
    const Params &P = c->P;
    if (P.problem != 6 && P.problem != 7 && P.problem != 8)
        return fail(c, GRHD_ERR_UNSUPPORTED,
                    "convergence norms exist for the advected wave, Michel and dust problems only");
    // row j = max(1, ny / 2), k = 1 of rho (output.f90:1166); one contiguous
    // run of the SoA density plane
    const int jr = std::max(1, P.ny / 2);
    const int js = (P.ny > 1) ? jr : 1, ks = 1;
    std::vector<double> rho((size_t)P.nx);
    HIPCHK(c, hipMemcpyAsync(rho.data(), c->p + pindex(P, 1, js, ks), rho.size() * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<double> exact((size_t)P.nx);
    int i0 = 1, i1 = P.nx;
    if (P.problem == 6) {
        // exact cell averages of the advected wave at time t (output.f90:1166-1185)
        const double pi = M_PI;
        const double arg = 0.5 * c->cfg.convergence_wave_number * pi * P.dx;
        const double sinc = (std::fabs(arg) > DBL_EPSILON) ? std::sin(arg) / arg : 1.0;
        for (int i = 1; i <= P.nx; i++) {
            const double xa = xc(P, i) - c->cfg.convergence_velocity * t;
            exact[(size_t)(i - 1)] = c->cfg.convergence_density +
                                     c->cfg.convergence_amplitude * sinc *
                                         std::sin(c->cfg.convergence_wave_number * pi * xa);
        }
    } else {
        // Michel (the initial data, output.f90:1186-1211) and dust (the
        // stationary free fall, 1212-1235): cells int(0.1 nx)..int(0.95 nx)
        std::vector<double> h;
        const int rc = initial_conditions(P, c->cfg, h, c->err);
        if (rc != GRHD_OK) return rc;
        const RefLayout L(P);
        for (int i = 1; i <= P.nx; i++) exact[(size_t)(i - 1)] = h[L.p(EQ_DE, i, jr, 1)];
        i0 = std::max(1, (int)(0.10 * (double)P.nx));
        i1 = std::min(P.nx, std::max(i0, (int)(0.95 * (double)P.nx)));
    }
    double s1 = 0.0, s2 = 0.0, linf = 0.0;
    int im = i0;
    for (int i = i0; i <= i1; i++) {
        const double diff = std::fabs(rho[(size_t)(i - 1)] - exact[(size_t)(i - 1)]);
        s1 = s1 + diff;
        s2 = s2 + diff * diff;
        if (diff > linf) {
            linf = diff;
            im = i;
        }
    }
    const double cells = (double)(i1 - i0 + 1);
    norms[0] = s1 / cells;
    norms[1] = std::sqrt(s2 / cells);
    norms[2] = linf;
    if (imax) *imax = im;
    return GRHD_OK;
}

int grhd_global_diagnostics(grhd_ctx *c, double *out, int64_t *counts)
{
    if (!c || !out || !counts) return GRHD_ERR_INVALID_ARGUMENT;
    const int nb = launch_global(c->P, c->stream, c->p, c->u[c->ucur], c->partial,
                                 (int)std::min<int64_t>(256, c->partial_cap / kGlobalPartials));
    HIPCHK(c, hipGetLastError());
    std::vector<double> h((size_t)nb * kGlobalPartials);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->partial, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double r[kGlobalPartials];
    for (int q = 0; q < kGlobalPartials; q++) r[q] = h[(size_t)q];
    for (int b = 1; b < nb; b++) {
        const double *x = &h[(size_t)b * kGlobalPartials];
        r[0] = r[0] + x[0];
        r[1] = r[1] + x[1];
        r[2] = r[2] + x[2];
        r[3] = std::fmin(r[3], x[3]);
        r[4] = std::fmin(r[4], x[4]);
        r[5] = std::fmax(r[5], x[5]);
        r[6] = std::fmax(r[6], x[6]);
        r[7] = std::fmax(r[7], x[7]);
        r[8] = r[8] + x[8];
        r[9] = r[9] + x[9];
    }
    // output.f90:1120-1122 order: mass, disk mass, angular momentum, rho min,
    // p min, rho max, p max, max v^2; atmosphere and invalid cells
    const double order[8] = {r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7]};
    for (int q = 0; q < 8; q++) out[q] = order[q];
    counts[0] = (int64_t)r[8];
    counts[1] = (int64_t)r[9];
    return GRHD_OK;
}

int grhd_ppi_modes(grhd_ctx *c, double *out)
{
    if (!c || !out) return GRHD_ERR_INVALID_ARGUMENT;
    const Params &P = c->P;
    if (P.nz < 2) return fail(c, GRHD_ERR_UNSUPPORTED, "azimuthal modes need nz >= 2");
    launch_plane_mass(P, c->stream, c->p, c->u[c->ucur], c->partial);
    HIPCHK(c, hipGetLastError());
    std::vector<double> m((size_t)P.nz);
    HIPCHK(c, hipMemcpyAsync(m.data(), c->partial, m.size() * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // the plane masses summed over k in the reference's order (output.f90:1021-1044)
    constexpr int M = kPpiModes;
    double c0 = 0.0, cr[M] = {0.0}, ci[M] = {0.0};
    for (int k = 1; k <= P.nz; k++) {
        const double mk = m[(size_t)(k - 1)];
        c0 = c0 + mk;
        for (int mode = 1; mode <= M; mode++) {
            cr[mode - 1] = cr[mode - 1] + mk * std::cos((double)mode * zc(P, k));
            ci[mode - 1] = ci[mode - 1] - mk * std::sin((double)mode * zc(P, k));
        }
    }
    out[0] = c0;
    for (int mode = 1; mode <= M; mode++) {
        double *o = out + 1 + 4 * (mode - 1);
        o[0] = cr[mode - 1];
        o[1] = ci[mode - 1];
        o[2] = 0.0;
        o[3] = 0.0;
        if (c0 > DBL_MIN) {
            o[2] = std::sqrt(cr[mode - 1] * cr[mode - 1] + ci[mode - 1] * ci[mode - 1]) / c0;
            o[3] = -std::atan2(ci[mode - 1], cr[mode - 1]) / (double)mode;
        }
    }
    return GRHD_OK;
}

int grhd_accretion_rate(grhd_ctx *c, double *mdot, int32_t *probe)
{
    if (!c || !mdot) return GRHD_ERR_INVALID_ARGUMENT;
    const Params &P = c->P;
    // the probe cell (initialization.f90:117-190): the first centre at or
    // outside the event horizon r+ = M + sqrt(M^2 - a^2) (a = 0 for EF)
    if (c->cfg.metric == GRHD_METRIC_MINKOWSKI || P.mass < 0.1)
        return fail(c, GRHD_ERR_UNSUPPORTED, "accretion rate needs a black-hole metric with M >= 0.1");
    const double a = (c->cfg.metric == GRHD_METRIC_KERR_SCHILD) ? P.spin : 0.0;
    const double rh = P.mass + std::sqrt(std::fmax(0.0, P.mass * P.mass - a * a));
    const double xh = (c->cfg.log_r != 0) ? std::log(rh) : rh;
    int ip = 0;
    if (xh < xf(P, 0)) {
        ip = 1;
    } else if (xh <= xf(P, P.nx)) {
        for (int i = 1; i <= P.nx; i++)
            if (xc(P, i) >= xh) {
                ip = i;
                break;
            }
    }
    if (ip == 0) return fail(c, GRHD_ERR_UNSUPPORTED, "no cell centre lies outside the event horizon");
    const int nb = launch_mdot(P, c->stream, ip, c->p, c->u[c->ucur], c->partial,
                               (int)std::min<int64_t>(256, c->partial_cap));
    HIPCHK(c, hipGetLastError());
    std::vector<double> h((size_t)nb);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->partial, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double s = 0.0;
    for (int b = 0; b < nb; b++) s = s + h[(size_t)b];
    *mdot = -s;
    if (probe) *probe = ip;
    return GRHD_OK;
}

int grhd_counters(grhd_ctx *c, int64_t *out)
{
    if (!c || !out) return GRHD_ERR_INVALID_ARGUMENT;
    unsigned long long h[4];
    HIPCHK(c, hipMemcpyAsync(h, c->counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out[0] = (int64_t)h[0];
    out[1] = (int64_t)h[1];
    out[2] = c->steps;
    return GRHD_OK;
}

int grhd_set_stream(grhd_ctx *c, void *s)
{
    if (!c) return GRHD_ERR_INVALID_ARGUMENT;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    c->stream = (hipStream_t)s;
    c->own_stream = false;
    return GRHD_OK;
}

int grhd_synchronize(grhd_ctx *c)
{
    if (!c) return GRHD_ERR_INVALID_ARGUMENT;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GRHD_OK;
}

int grhd_set_stage_form(grhd_ctx *c, int32_t zseg)
{
    if (!c || zseg < kStageXFused) return GRHD_ERR_INVALID_ARGUMENT;
    if ((zseg > 0 || zseg == kStageXFused) && (!zmarch_possible(c->P) || !c->geo[0] || !c->geo[2]))
        return fail(c, GRHD_ERR_UNSUPPORTED, "the z-march and x-fused forms need a 3D grid above 65,536 cells");
    c->forced_zseg = zseg;
    c->faces_tuned = false;  // applied (or timed) before the next step
    return GRHD_OK;
}

int grhd_set_face_form(grhd_ctx *c, int32_t form)
{
    if (!c || !(form == -1 || form == 1 || form == 2 || form == 4 || form == 8 || face_form_table(form)))
        return GRHD_ERR_INVALID_ARGUMENT;
    if (face_form_table(form) && !c->geo[0])
        return fail(c, GRHD_ERR_UNSUPPORTED, "the table forms need a grid outside the fused stage form");
    if (face_form_defers(form) && !c->defer_list[0])
        return fail(c, GRHD_ERR_UNSUPPORTED, "deferred faces need the compiled MP5 + sensor + HLLC scheme");
    c->forced_face = form;
    c->faces_tuned = false;  // applied before the next step
    return GRHD_OK;
}

int grhd_last_timing(grhd_ctx *c, double *ms, int64_t *launches)
{
    if (!c) return GRHD_ERR_INVALID_ARGUMENT;
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float f = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&f, c->ev0, c->ev1));
    if (ms) *ms = f;
    if (launches) *launches = c->last_launches;
    return GRHD_OK;
}

}  // extern "C"

// ----------------------------------------------------------- no allocation --
// grhd_step_noalloc keeps every intermediate of the step -- face fluxes, the
// RK stage buffer, the Michel injector, the coordinate table and the
// counters -- in one file-scope device array.  Each device that loads the
// library holds its own copy for the life of the process, at its own
// address: in host code `scratch` names only a host-side stand-in, so the
// address of the current device's copy comes from hipGetSymbolAddress, once
// per device, and only that address reaches kernels and copies.
namespace {

constexpr size_t kScratchBytes = (size_t)GRHD_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[kMaxDevices] = {};

// Address of the current device's scratch array (looked up on the first call
// on that device).
int scratch_address(unsigned char **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(nullptr, GRHD_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(nullptr, GRHD_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(nullptr, GRHD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch_dev[dev];
    return GRHD_OK;
}

__global__ void k_clear_counters(unsigned long long *c)
{
    if (threadIdx.x < 4) c[threadIdx.x] = 0ull;
}

// The Michel outer injector (conditions.f90:229-246) from the outer radial
// ghost layers of p, which hold it: the initial data puts it there and every
// ghost fill writes it back.  injector[(((k-1) ny + j-1) ng + g-1) 5 + v].
__global__ void k_injector_from_ghosts(Params P, const double *__restrict__ p, double *__restrict__ inj)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)P.ng * P.ny * P.nz) return;
    const int g = (int)(t % P.ng) + 1;
    const int64_t r = t / P.ng;
    const int j = (int)(r % P.ny) + 1, k = (int)(r / P.ny) + 1;
    const int64_t o = pindex(P, P.nx + g, j, k);
    for (int v = 0; v < NEQ; v++) inj[t * NEQ + v] = p[v * P.vstride + o];
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Uploads P's coordinate table to dst through a fixed host staging buffer
// (no allocation); synchronous on s.
std::mutex g_stage_mu;
double g_stage[4096];

int upload_coord(const Params &P, double *dst, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_stage_mu);
    size_t n = 0, done = 0;
    hipError_t e = hipSuccess;
    auto flush = [&]() {
        if (n == 0 || e != hipSuccess) return;
        e = hipMemcpyAsync(dst + done, g_stage, n * sizeof(double), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        done += n;
        n = 0;
    };
    coord_entries(P, [&](double v) {
        g_stage[n++] = v;
        if (n == sizeof(g_stage) / sizeof(g_stage[0])) flush();
    });
    flush();
    if (e != hipSuccess) return fail(nullptr, GRHD_ERR_HIP, std::string("coordinate table upload: ") + hipGetErrorString(e));
    return GRHD_OK;
}

}  // namespace

extern "C" int grhd_step_noalloc(const grhd_config *cfg, double *p, double *u, double dt, int32_t nsteps,
                                 void *stream)
{
    if (!cfg || !p || !u || nsteps < 0) return fail(nullptr, GRHD_ERR_INVALID_ARGUMENT, "null argument or nsteps < 0");
    std::string msg;
    if (validate(*cfg, msg)) return fail(nullptr, GRHD_ERR_INVALID_CONFIG, msg);
    grhd_config c = *cfg;
    if (c.metric == GRHD_METRIC_KERR_SCHILD) c.geometry = GRHD_GEOM_SPHEROIDAL;
    Params P = make_params(c);
    if ((P.problem == 5 && (P.nx < P.ng || (P.nz > 1 && P.nz < P.ng))) || (P.problem == 11 && P.ny < P.ng) ||
        (P.problem == 12 && P.ny < P.ng))
        return fail(nullptr, GRHD_ERR_INVALID_CONFIG, "grid smaller than the ghost depth");
    const bool mixed = (cfg->precision == GRHD_PRECISION_MIXED);
    const size_t rsz = mixed ? sizeof(float) : sizeof(double);
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    const int64_t fstr[3] = {(int64_t)(P.nx + 1) * P.ny * P.nz, (int64_t)P.nx * (P.ny + 1) * P.nz,
                             (int64_t)P.nx * P.ny * (P.nz + 1)};
    // carve the array: faces of the active directions, the stage buffer, the
    // injector, the coordinate table, the counters
    size_t off[7], need = 0;
    for (int d = 0; d < 3; d++) {
        const bool active = (d == 0) || (d == 1 && P.ny > 1) || (d == 2 && P.nz > 1);
        off[d] = need;
        need += align256((active ? (size_t)fstr[d] * NEQ : 1) * rsz);
    }
    off[3] = need;
    need += align256((size_t)ncell * NEQ * sizeof(double));
    off[4] = need;
    need += align256((size_t)NEQ * P.ng * P.ny * P.nz * sizeof(double) + 8);
    off[5] = need;
    need += align256((size_t)coord_table_size(P.nx, P.ny, P.ng) * sizeof(double));
    off[6] = need;
    need += align256(4 * sizeof(unsigned long long));
    if (need > kScratchBytes)
        return fail(nullptr, GRHD_ERR_CAPACITY,
                    "grhd_step_noalloc: the problem needs " + std::to_string(need) +
                        " bytes of intermediates, the scratch array holds " + std::to_string(kScratchBytes));
    unsigned char *base = nullptr;
    int rc = scratch_address(&base);
    if (rc != GRHD_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    void *F[3] = {base + off[0], base + off[1], base + off[2]};
    double *ustage = reinterpret_cast<double *>(base + off[3]);
    double *inj = reinterpret_cast<double *>(base + off[4]);
    double *coord = reinterpret_cast<double *>(base + off[5]);
    unsigned long long *counters = reinterpret_cast<unsigned long long *>(base + off[6]);
    rc = upload_coord(P, coord, s);
    if (rc != GRHD_OK) return rc;
    P.coord = coord;
    // the array keeps the last call's contents: the counters start at zero
    k_clear_counters<<<1, 64, 0, s>>>(counters);
    if (P.problem == 7)
        k_injector_from_ghosts<<<(unsigned)(((int64_t)P.ng * P.ny * P.nz + 255) / 256), 256, 0, s>>>(P, p, inj);
    launch_boundaries(P, s, p, inj);  // grhd2.f90:233
    double *ua = u, *ub = ustage;
    for (int st = 0; st < nsteps; st++) {
        for (int rk = 1; rk <= 3; rk++) launch_stage(P, s, p, F, fstr, ua, ub, rk, dt, counters, mixed, inj, true);
        std::swap(ua, ub);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && ua != u)
        e = hipMemcpyAsync(u, ua, (size_t)ncell * NEQ * sizeof(double), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return fail(nullptr, GRHD_ERR_HIP, std::string("grhd_step_noalloc: ") + hipGetErrorString(e));
    return GRHD_OK;
}
