// grhd_kernels.hip -- device kernels of the SSP-RK3 finite-volume update.
//
// One RK stage = face-flux kernels (one per active direction) + one cell
// kernel (flux divergence, sources, RK combination, con2prim) + one boundary
// kernel.  Kernels are templated on the metric kind so each launch evaluates
// its analytic metric with compile-time structure; reconstruction and the
// Riemann solver are uniform runtime branches.  This unit instantiates the FP64
// stage (compiled without FMA contraction) and the boundary, conversion and CFL
// kernels.
#include "grhd_stage.hpp"

namespace grhd {

// prim_to_cons over the interior (grhd2.f90:126-134).
template <int MK>
__global__ __launch_bounds__(256) void k_p2c(Params P, const double *__restrict__ p,
                                             double *__restrict__ u)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    if (t >= ncell) return;
    const int i = (int)(t % P.nx) + 1;
    const int64_t rem = t / P.nx;
    const int j = (int)(rem % P.ny) + 1, k = (int)(rem / P.ny) + 1;
    double g[3][3], gi[3][3], sqg, alpha, beta[3], q[NEQ], U[NEQ];
    metric_c2p<MK>(P, xc(P, i), yc(P, j), g, gi, sqg, alpha, beta);
    const int64_t pc = pindex(P, i, j, k);
    for (int v = 0; v < NEQ; v++) q[v] = p[v * P.vstride + pc];
    p2c<double>(P.g1, q, Mat3<double>{g}, sqg, U);
    for (int v = 0; v < NEQ; v++) u[v * ncell + t] = U[v];
}

// cons_to_prim over the interior; U repaired in place.
template <int MK>
__global__ __launch_bounds__(256) void k_c2p(Params P, double *__restrict__ p, double *__restrict__ u,
                                             unsigned long long *__restrict__ counters)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    if (t >= ncell) return;
    const int i = (int)(t % P.nx) + 1;
    const int64_t rem = t / P.nx;
    const int j = (int)(rem % P.ny) + 1, k = (int)(rem / P.ny) + 1;
    double g[3][3], gi[3][3], sqg, alpha, beta[3], q[NEQ], c[NEQ];
    metric_c2p<MK>(P, xc(P, i), yc(P, j), g, gi, sqg, alpha, beta);
    for (int v = 0; v < NEQ; v++) c[v] = u[v * ncell + t];
    int brent = 0;
    if (c2p(P.g1, P.rho_floor, P.p_floor, g, gi, sqg, c, q, brent)) atomicAdd(&counters[0], 1ull);
    if (brent) atomicAdd(&counters[1], 1ull);
    const int64_t pc = pindex(P, i, j, k);
    for (int v = 0; v < NEQ; v++) {
        u[v * ncell + t] = c[v];
        p[v * P.vstride + pc] = q[v];
    }
}

// ------------------------------------------------------------- boundaries --
// set_boundary_conditions (conditions.f90:773-1066) restricted to the ghost
// cells the update reads: for each active axis, the ghost layers of the rows
// whose transverse indices are interior.  Every rule there is a gather from
// interior cells (or analytic / cached boundary data), so each ghost cell is
// filled independently.  The outermost layer (index -ng) is never written,
// as in the reference (it keeps its initial value).
GRHD_HD int wrap(int m, int n) { return ((m - 1) % n + n) % n + 1; }

template <int MK>
__global__ __launch_bounds__(256) void k_boundaries(Params P, double *__restrict__ p,
                                                    const double *__restrict__ injector)
{
    const int ng = P.ng, nx = P.nx, ny = P.ny, nz = P.nz;
    const int64_t nxs = 2LL * ng * ny * nz;
    const int64_t nys = (ny > 1) ? 2LL * ng * nx * nz : 0;
    const int64_t nzs = (nz > 1) ? 2LL * ng * nx * ny : 0;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nxs + nys + nzs) return;
    const int prob = P.problem;
    double q[NEQ];
    int gi, gj, gk;  // ghost cell (reference indices)
    auto load = [&](int i, int j, int k) {
        const int64_t o = pindex(P, i, j, k);
        for (int v = 0; v < NEQ; v++) q[v] = p[v * P.vstride + o];
    };
    if (t < nxs) {  // ---- radial / x ghosts at interior (j, k)
        const int g = (int)(t % ng) + 1;
        int64_t r = t / ng;
        const int hi = (int)(r % 2);
        r /= 2;
        gj = (int)(r % ny) + 1;
        gk = (int)(r / ny) + 1;
        gi = hi ? nx + g : 1 - g;
        switch (prob) {
        case 1: case 2: case 3:  // outflow copy (conditions.f90:795-798)
            load(hi ? nx : 1, gj, gk);
            break;
        case 4: case 6:  // periodic (conditions.f90:809-811)
            load(wrap(gi, nx), gj, gk);
            break;
        case 5:  // jet: reflective low, ambient high (conditions.f90:822-847)
            if (hi) {
                q[EQ_DE] = P.jet_ambient_density;
                q[EQ_PR] = P.jet_ambient_pressure;
                q[EQ_VX] = 0.0;
                q[EQ_VY] = 0.0;
                q[EQ_VZ] = P.jet_ambient_velocity;
            } else {
                load(g, gj, gk);
                q[EQ_VX] = -q[EQ_VX];
            }
            break;
        case 7:  // inner copy, outer analytic injector (conditions.f90:873-883)
            if (hi) {
                for (int v = 0; v < NEQ; v++)
                    q[v] = injector[((((int64_t)(gk - 1) * ny + (gj - 1)) * ng + (g - 1)) * NEQ) + v];
            } else {
                load(1, gj, gk);
            }
            break;
        case 8:  // inner outflow towards the hole, outer dust injector (conditions.f90:896-907)
            if (hi) {
                double r_loc, jac;
                const double xr = xc(P, nx + g);
                if (MK == MK_EF_LOG || MK == MK_KS_LOG || MK == MK_MINK_SPH_LOG) {
                    r_loc = exp(xr);
                    jac = r_loc;
                } else {
                    r_loc = xr;
                    jac = 1.0;
                }
                const double M = P.mass, ff = sqrt(2.0 * M / r_loc);
                q[EQ_DE] = -P.dust_accretion_constant / ((r_loc * r_loc) * ff);
                q[EQ_PR] = P.p_floor;
                q[EQ_VX] = -1.0 / (sqrt(1.0 + r_loc / (2.0 * M)) * (1.0 + ff + 2.0 * M / r_loc));
                q[EQ_VX] = q[EQ_VX] / jac;
                q[EQ_VY] = 0.0;
                q[EQ_VZ] = 0.0;
            } else {
                load(1, gj, gk);
                q[EQ_DE] = rmax(q[EQ_DE], P.rho_floor);
                q[EQ_PR] = rmax(q[EQ_PR], P.p_floor);
                q[EQ_VX] = rmin(0.0, q[EQ_VX]);
            }
            break;
        default: {  // 9-12: shift-aware radial outflow (conditions.f90:918-938, 951-973, 1007-1026)
            FaceMetric m;
            const int ic = hi ? nx : 1;
            metric_face<MK>(P, xc(P, ic), yc(P, gj), m);
            const double vlim = m.beta[0] / m.alpha;
            load(ic, gj, gk);
            const double rho_edge = q[EQ_DE];
            q[EQ_DE] = rmax(q[EQ_DE], P.rho_floor);
            q[EQ_PR] = rmax(q[EQ_PR], P.p_floor);
            if (hi) {
                q[EQ_VX] = rmax(vlim, q[EQ_VX]);
            } else {
                q[EQ_VX] = rmin(vlim, q[EQ_VX]);
                if (prob <= 10 && rho_edge < 1.0e-6) {
                    q[EQ_VY] = 0.0;
                    q[EQ_VZ] = 0.0;
                }
            }
            break;
        }
        }
    } else if (t < nxs + nys) {  // ---- polar / y ghosts at interior (i, k)
        t -= nxs;
        const int g = (int)(t % ng) + 1;
        int64_t r = t / ng;
        const int hi = (int)(r % 2);
        r /= 2;
        gi = (int)(r % nx) + 1;
        gk = (int)(r / nx) + 1;
        gj = hi ? ny + g : 1 - g;
        if (prob == 5) {  // jet: ghosts stay ambient (conditions.f90:824-835)
            q[EQ_DE] = P.jet_ambient_density;
            q[EQ_PR] = P.jet_ambient_pressure;
            q[EQ_VX] = 0.0;
            q[EQ_VY] = 0.0;
            q[EQ_VZ] = P.jet_ambient_velocity;
        } else if (prob == 11) {  // reflective theta (conditions.f90:977-987)
            load(gi, hi ? ny + 1 - g : g, gk);
            q[EQ_VY] = -q[EQ_VY];
        } else if (prob == 12) {  // trans-polar with phi + pi (conditions.f90:1032-1050)
            int ks = gk + nz / 2;
            if (ks > nz) ks -= nz;
            load(gi, hi ? ny + 1 - g : g, ks);
            q[EQ_VY] = -q[EQ_VY];
        } else {  // periodic
            load(gi, wrap(gj, ny), gk);
        }
    } else {  // ---- azimuthal / z ghosts at interior (i, j)
        t -= nxs + nys;
        const int g = (int)(t % ng) + 1;
        int64_t r = t / ng;
        const int hi = (int)(r % 2);
        r /= 2;
        gi = (int)(r % nx) + 1;
        gj = (int)(r / nx) + 1;
        gk = hi ? nz + g : 1 - g;
        if (prob == 5) {  // jet nozzle and outflow (conditions.f90:851-869)
            if (hi) {
                load(gi, gj, nz);
                q[EQ_VZ] = rmax(0.0, q[EQ_VZ]);
            } else if (xc(P, gi) <= P.jet_nozzle_radius) {
                q[EQ_DE] = P.jet_density;
                q[EQ_PR] = P.jet_pressure;
                q[EQ_VX] = 0.0;
                q[EQ_VY] = 0.0;
                q[EQ_VZ] = P.jet_velocity;
            } else {
                load(gi, gj, g);
                q[EQ_VZ] = -q[EQ_VZ];
            }
        } else {
            load(gi, gj, wrap(gk, nz));
        }
    }
    const int64_t o = pindex(P, gi, gj, gk);
    for (int v = 0; v < NEQ; v++) p[v * P.vstride + o] = q[v];
    // The jet fill first resets every ghost cell to the ambient state
    // (conditions.f90:824-835), the outermost layer -ng included; the other
    // problems never write that layer.
    if (prob == 5 && (gi == 1 - ng || gj == 1 - ng || gk == 1 - ng)) {
        const int64_t oo = pindex(P, gi == 1 - ng ? -ng : gi, gj == 1 - ng ? -ng : gj,
                                  gk == 1 - ng ? -ng : gk);
        p[0 * P.vstride + oo] = P.jet_ambient_density;
        p[1 * P.vstride + oo] = P.jet_ambient_pressure;
        p[2 * P.vstride + oo] = 0.0;
        p[3 * P.vstride + oo] = 0.0;
        p[4 * P.vstride + oo] = P.jet_ambient_velocity;
    }
}

// ------------------------------------------------------------------ CFL --
// update_adaptive_dt (evolution.f90:218-255): per-(i, j) light-cone speeds
// (the metric does not depend on the third coordinate), wave64 max-reduction,
// one partial per block.
template <int MK>
__global__ __launch_bounds__(256) void k_cfl(Params P, double *__restrict__ partial)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double vx = 0.0, vy = 0.0, vz = 0.0, bad = 0.0;
    if (t < (int64_t)P.nx * P.ny) {
        const int i = (int)(t % P.nx) + 1, j = (int)(t / P.nx) + 1;
        CentreMetric m;
        metric_centre<MK>(P, xc(P, i), yc(P, j), m);
        const double a = m.alpha;
        if (!isfinite(a) || a <= 0.0 || !isfinite(m.beta[0]) || !isfinite(m.beta[1]) ||
            !isfinite(m.beta[2]) || !isfinite(m.ginv[0][0]) || m.ginv[0][0] <= 0.0 ||
            (P.ny > 1 && (!isfinite(m.ginv[1][1]) || m.ginv[1][1] <= 0.0)) ||
            (P.nz > 1 && (!isfinite(m.ginv[2][2]) || m.ginv[2][2] <= 0.0))) {
            bad = 1.0;
        } else {
            vx = a * sqrt(m.ginv[0][0]) + fabs(m.beta[0]);
            if (P.ny > 1) vy = a * sqrt(m.ginv[1][1]) + fabs(m.beta[1]);
            if (P.nz > 1) vz = a * sqrt(m.ginv[2][2]) + fabs(m.beta[2]);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        vx = rmax(vx, __shfl_xor(vx, off, 64));
        vy = rmax(vy, __shfl_xor(vy, off, 64));
        vz = rmax(vz, __shfl_xor(vz, off, 64));
        bad = bad + __shfl_xor(bad, off, 64);
    }
    __shared__ double red[4][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[wave][0] = vx;
        red[wave][1] = vy;
        red[wave][2] = vz;
        red[wave][3] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) / 64;
        double r0 = red[0][0], r1 = red[0][1], r2 = red[0][2], r3 = red[0][3];
        for (int w = 1; w < nw; w++) {
            r0 = rmax(r0, red[w][0]);
            r1 = rmax(r1, red[w][1]);
            r2 = rmax(r2, red[w][2]);
            r3 = r3 + red[w][3];
        }
        partial[4 * blockIdx.x + 0] = r0;
        partial[4 * blockIdx.x + 1] = r1;
        partial[4 * blockIdx.x + 2] = r2;
        partial[4 * blockIdx.x + 3] = r3;
    }
}

// --------------------------------------------------------------- launchers --

template <int MK>
void launch_boundaries_mk(const Params &P, hipStream_t s, double *p, const double *injector)
{
    const int64_t n = 2LL * P.ng * P.ny * P.nz + (P.ny > 1 ? 2LL * P.ng * P.nx * P.nz : 0) +
                      (P.nz > 1 ? 2LL * P.ng * P.nx * P.ny : 0);
    k_boundaries<MK><<<blocks_for(n, 256), 256, 0, s>>>(P, p, injector);
}

#define GRHD_DISPATCH(MKV, CALL)                                      \
    switch (MKV) {                                                    \
    case MK_MINK_CART: { constexpr int MKC = MK_MINK_CART; CALL; } break;       \
    case MK_MINK_CYL: { constexpr int MKC = MK_MINK_CYL; CALL; } break;         \
    case MK_MINK_SPH: { constexpr int MKC = MK_MINK_SPH; CALL; } break;         \
    case MK_MINK_SPH_LOG: { constexpr int MKC = MK_MINK_SPH_LOG; CALL; } break; \
    case MK_EF: { constexpr int MKC = MK_EF; CALL; } break;                     \
    case MK_EF_LOG: { constexpr int MKC = MK_EF_LOG; CALL; } break;             \
    case MK_KS: { constexpr int MKC = MK_KS; CALL; } break;                     \
    default: { constexpr int MKC = MK_KS_LOG; CALL; } break;                    \
    }

void launch_stage_f64(const Params &P, hipStream_t s, double *p, void *const F[3],
                      const int64_t fstr[3], const double *u0, double *u, int stage, double dt,
                      unsigned long long *counters)
{
    GRHD_STAGE_DISPATCH(double)
}

void launch_stage(const Params &P, hipStream_t s, double *p, void *const F[3], const int64_t fstr[3],
                  const double *u0, double *u, int stage, double dt, unsigned long long *counters,
                  bool mixed, const double *injector, bool with_boundaries)
{
    if (mixed) launch_stage_f32(P, s, p, F, fstr, u0, u, stage, dt, counters);
    else launch_stage_f64(P, s, p, F, fstr, u0, u, stage, dt, counters);
    if (with_boundaries) launch_boundaries(P, s, p, injector);
}

void launch_boundaries(const Params &P, hipStream_t s, double *p, const double *injector)
{
    GRHD_DISPATCH(P.mk, (launch_boundaries_mk<MKC>(P, s, p, injector)));
}

void launch_p2c(const Params &P, hipStream_t s, const double *p, double *u)
{
    const int64_t n = (int64_t)P.nx * P.ny * P.nz;
    GRHD_DISPATCH(P.mk, (k_p2c<MKC><<<blocks_for(n, 256), 256, 0, s>>>(P, p, u)));
}

void launch_c2p(const Params &P, hipStream_t s, double *p, double *u, unsigned long long *counters)
{
    const int64_t n = (int64_t)P.nx * P.ny * P.nz;
    GRHD_DISPATCH(P.mk, (k_c2p<MKC><<<blocks_for(n, 256), 256, 0, s>>>(P, p, u, counters)));
}

int launch_cfl(const Params &P, hipStream_t s, double *partial)
{
    const int64_t n = (int64_t)P.nx * P.ny;
    const unsigned nb = blocks_for(n, 256);
    GRHD_DISPATCH(P.mk, (k_cfl<MKC><<<nb, 256, 0, s>>>(P, partial)));
    return (int)nb;
}

}  // namespace grhd
