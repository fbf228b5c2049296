// grhd_kernels.hip -- device kernels of the SSP-RK3 finite-volume update.
//
// One RK stage = face-flux kernels (one per active direction) + one cell
// kernel (flux divergence, sources, RK combination, con2prim) + one boundary
// kernel.  Kernels are templated on the metric kind so each launch evaluates
// its analytic metric with compile-time structure; reconstruction and the
// Riemann solver are uniform runtime branches.
#include "grhd_kernels.hpp"

namespace grhd {

template <int MK> constexpr bool kSRHD = (MK == MK_MINK_CART);

// -------------------------------------------------------------- fluxes --
// Face arrays have shape [5][nfz][nfy][nfx] with nf_d = n_d + (d == DIR).
// The analytic metric depends on (r, theta) only, so each thread owns a run
// of P.kc consecutive phi indices at fixed (i, j): the face metric is
// evaluated once and reused along the run.  Lanes of a wave are consecutive
// in i, so every load and store of the run is coalesced.
template <int MK, int DIR, typename R>
__device__ __forceinline__ void flux_face(const Params &P, const double *__restrict__ p,
                                          const FaceR<R> &f, int i, int j, int k, int fi, int nodes,
                                          R out[NEQ])
{
    R s[NEQ][6];
    const int lo = (P.ng >= 2) ? -2 : 0, hi = (P.ng >= 3) ? 3 : (P.ng == 2 ? 2 : 1);
    const int64_t step = (DIR == 1) ? 1 : (DIR == 2) ? (int64_t)P.NXs : (int64_t)P.NXs * P.NYs;
    const int64_t base = pindex(P, DIR == 1 ? fi : i, DIR == 2 ? fi : j, DIR == 3 ? fi : k);
#pragma unroll
    for (int o = -2; o <= 3; o++) {
        const bool in = (o >= lo && o <= hi);
#pragma unroll
        for (int v = 0; v < NEQ; v++) s[v][o + 2] = in ? (R)p[v * P.vstride + base + o * step] : (R)0;
    }
    R qL[NEQ], qR[NEQ];
    recon_face<R>(P, s, fi, nodes, DIR, qL, qR);
    riemann<R, DIR, kSRHD<MK>>(P.rs, (R)P.g1, (R)P.gam, qL, qR, f, out);
}

template <int MK, int DIR, typename R>
__global__ __launch_bounds__(256) void k_flux(Params P, const double *__restrict__ p,
                                              R *__restrict__ F, int64_t fstride)
{
    const int nfx = P.nx + (DIR == 1), nfy = P.ny + (DIR == 2), nfz = P.nz + (DIR == 3);
    const int nchunk = (nfz + P.kc - 1) / P.kc;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nfx * nfy * nchunk) return;
    const int a = (int)(t % nfx);
    const int64_t rem = t / nfx;
    const int b = (int)(rem % nfy), chunk = (int)(rem / nfy);
    const int c0 = chunk * P.kc, c1 = min(c0 + P.kc, nfz);
    // reference indices: faces 0..n along DIR, cells 1..n across
    const int i = a + (DIR != 1), j = b + (DIR != 2);
    const int nodes = (DIR == 1) ? P.nx : (DIR == 2) ? P.ny : P.nz;
    const double eps64 = 64.0 * DBL_EPSILON;
    bool zero = false;
    if (DIR == 1 && P.cyl) zero = fabs(xf(P, i)) <= eps64;           // evolution.f90:69-71
    if (DIR == 2 && P.polar) zero = fabs(sin(yf(P, j))) <= eps64;    // evolution.f90:122-124
    FaceMetric m;
    if (DIR == 1) metric_face<MK>(P, xf(P, i), yc(P, j), m);
    else if (DIR == 2) metric_face<MK>(P, xc(P, i), yf(P, j), m);
    else metric_face<MK>(P, xc(P, i), yc(P, j), m);  // z faces: centre metric
    FaceR<R> f;
    f.alpha = (R)m.alpha;
    f.sqg = (R)sqrt(m.det);
    for (int x = 0; x < 3; x++) {
        f.beta[x] = (R)m.beta[x];
        for (int y = 0; y < 3; y++) f.g[x][y] = (R)m.g[x][y];
    }
    for (int c = c0; c < c1; c++) {
        const int k = c + (DIR != 3);
        const int fi = (DIR == 1) ? i : (DIR == 2) ? j : k;
        R out[NEQ];
        if (zero) {
#pragma unroll
            for (int v = 0; v < NEQ; v++) out[v] = 0;
        } else {
            flux_face<MK, DIR, R>(P, p, f, i, j, k, fi, nodes, out);
        }
        const int64_t fo = ((int64_t)c * nfy + b) * nfx + a;
#pragma unroll
        for (int v = 0; v < NEQ; v++) F[v * fstride + fo] = out[v];
    }
}

// --------------------------------------------------------- cell update --
// stage 1..3: rhs (evolution.f90:85-88,137-139,180-182), RK combination
// (grhd2.f90:243-249), cons_to_prim (grhd2.f90:252-260).  stage 0 writes the
// rhs itself to u (parity entry point).  Each thread owns a run of P.kc cells
// along phi and evaluates the centre metric (Christoffels, dg, gamma^ij)
// once for the run.
template <int MK, typename R>
__global__ __launch_bounds__(256) void k_update(Params P, double *__restrict__ p,
                                                const R *__restrict__ Fx, const R *__restrict__ Fy,
                                                const R *__restrict__ Fz, int64_t fsx, int64_t fsy,
                                                int64_t fsz, const double *__restrict__ u0,
                                                double *__restrict__ u, int stage, double dt,
                                                unsigned long long *__restrict__ counters)
{
    const int nchunk = (P.nz + P.kc - 1) / P.kc;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)P.nx * P.ny * nchunk) return;
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    const int i0 = (int)(t % P.nx);
    const int64_t rem = t / P.nx;
    const int j0 = (int)(rem % P.ny), chunk = (int)(rem / P.ny);
    const int i = i0 + 1, j = j0 + 1;
    CentreMetric m;
    metric_centre<MK>(P, xc(P, i), yc(P, j), m);
    const int k0b = chunk * P.kc, k0e = min(k0b + P.kc, P.nz);
    unsigned n_atm = 0, n_brent = 0;
    for (int k0 = k0b; k0 < k0e; k0++) {
        const int k = k0 + 1;
        const int64_t tc = ((int64_t)k0 * P.ny + j0) * P.nx + i0;
        const int64_t pc = pindex(P, i, j, k);
        double q[NEQ], rhs[NEQ], src[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) q[v] = p[v * P.vstride + pc];
        if (MK != MK_MINK_CART) {
            sources(P.g1, m, q, src);
        } else {
#pragma unroll
            for (int v = 0; v < NEQ; v++) src[v] = 0.0;
        }
        {
            const int64_t fa = ((int64_t)k0 * P.ny + j0) * (P.nx + 1) + i0;
#pragma unroll
            for (int v = 0; v < NEQ; v++)
                rhs[v] = -((double)Fx[v * fsx + fa + 1] - (double)Fx[v * fsx + fa]) / P.dx + src[v];
        }
        if (P.ny > 1) {
            const int64_t fa = ((int64_t)k0 * (P.ny + 1) + j0) * P.nx + i0;
#pragma unroll
            for (int v = 0; v < NEQ; v++)
                rhs[v] = rhs[v] - ((double)Fy[v * fsy + fa + P.nx] - (double)Fy[v * fsy + fa]) / P.dy;
        }
        if (P.nz > 1) {
            const int64_t plane = (int64_t)P.nx * P.ny;
#pragma unroll
            for (int v = 0; v < NEQ; v++)
                rhs[v] = rhs[v] - ((double)Fz[v * fsz + tc + plane] - (double)Fz[v * fsz + tc]) / P.dz;
        }
        if (stage == 0) {
#pragma unroll
            for (int v = 0; v < NEQ; v++) u[v * ncell + tc] = rhs[v];
            continue;
        }
        double c[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
            const double a0 = u0[v * ncell + tc];
            if (stage == 1) c[v] = a0 + rhs[v] * dt;
            else if (stage == 2) c[v] = 0.75 * a0 + 0.25 * (u[v * ncell + tc] + rhs[v] * dt);
            else c[v] = (a0 + 2.0 * (u[v * ncell + tc] + rhs[v] * dt)) / 3.0;
        }
        int brent = 0;
        n_atm += c2p(P.g1, P.rho_floor, P.p_floor, m.g, m.ginv, m.sqg, c, q, brent) ? 1u : 0u;
        n_brent += brent ? 1u : 0u;
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
            u[v * ncell + tc] = c[v];
            p[v * P.vstride + pc] = q[v];
        }
    }
    if (n_atm) atomicAdd(&counters[0], (unsigned long long)n_atm);
    if (n_brent) atomicAdd(&counters[1], (unsigned long long)n_brent);
}

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
    p2c<double>(P.g1, q, g, sqg, U);
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
static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <int MK>
static void launch_stage_mk(const Params &P, hipStream_t s, double *p, void *const F[3],
                            const int64_t fstr[3], const double *u0, double *u, int stage, double dt,
                            unsigned long long *counters, bool mixed, const double *injector,
                            bool with_boundaries)
{
    const int bs = 256;
    const int kc = P.kc;
    auto chunks = [kc](int n) { return (int64_t)((n + kc - 1) / kc); };
    const int64_t nfx = (int64_t)(P.nx + 1) * P.ny * chunks(P.nz);
    const int64_t nfy = (int64_t)P.nx * (P.ny + 1) * chunks(P.nz);
    const int64_t nfz = (int64_t)P.nx * P.ny * chunks(P.nz + 1);
    const int64_t ncol = (int64_t)P.nx * P.ny * chunks(P.nz);
    if (!mixed) {
        double *Fx = (double *)F[0], *Fy = (double *)F[1], *Fz = (double *)F[2];
        k_flux<MK, 1, double><<<blocks_for(nfx, bs), bs, 0, s>>>(P, p, Fx, fstr[0]);
        if (P.ny > 1) k_flux<MK, 2, double><<<blocks_for(nfy, bs), bs, 0, s>>>(P, p, Fy, fstr[1]);
        if (P.nz > 1) k_flux<MK, 3, double><<<blocks_for(nfz, bs), bs, 0, s>>>(P, p, Fz, fstr[2]);
        k_update<MK, double><<<blocks_for(ncol, bs), bs, 0, s>>>(P, p, Fx, Fy, Fz, fstr[0], fstr[1],
                                                                 fstr[2], u0, u, stage, dt, counters);
    } else {
        float *Fx = (float *)F[0], *Fy = (float *)F[1], *Fz = (float *)F[2];
        k_flux<MK, 1, float><<<blocks_for(nfx, bs), bs, 0, s>>>(P, p, Fx, fstr[0]);
        if (P.ny > 1) k_flux<MK, 2, float><<<blocks_for(nfy, bs), bs, 0, s>>>(P, p, Fy, fstr[1]);
        if (P.nz > 1) k_flux<MK, 3, float><<<blocks_for(nfz, bs), bs, 0, s>>>(P, p, Fz, fstr[2]);
        k_update<MK, float><<<blocks_for(ncol, bs), bs, 0, s>>>(P, p, Fx, Fy, Fz, fstr[0], fstr[1],
                                                                fstr[2], u0, u, stage, dt, counters);
    }
    if (with_boundaries) launch_boundaries_mk<MK>(P, s, p, injector);
}

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

void launch_stage(const Params &P, hipStream_t s, double *p, void *const F[3], const int64_t fstr[3],
                  const double *u0, double *u, int stage, double dt, unsigned long long *counters,
                  bool mixed, const double *injector, bool with_boundaries)
{
    GRHD_DISPATCH(P.mk, (launch_stage_mk<MKC>(P, s, p, F, fstr, u0, u, stage, dt, counters, mixed,
                                              injector, with_boundaries)));
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
