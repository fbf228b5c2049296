// grhd_stage.hpp -- templated kernels of one RK stage (face fluxes and the
// cell update) and their launcher.  Included by the two translation units that
// instantiate them: grhd_kernels.hip (FP64 faces, reference rounding sequence)
// and grhd_kernels_mixed.hip (FP32 faces).
#pragma once
#include "grhd_kernels.hpp"

namespace grhd {

static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <int MK> constexpr bool kSRHD = (MK == MK_MINK_CART);

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8), and each XCD has its own L2.  Renumbering so that every XCD walks a
// contiguous range of logical blocks keeps neighbouring rows -- which share
// stencil data -- in one L2.  A bijection of [0, nb) for any nb.
constexpr unsigned kXCD = 8;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb)
{
    const unsigned per = nb / kXCD, rem = nb % kXCD, x = b % kXCD, q = b / kXCD;
    return x * per + (x < rem ? x : rem) + q;
}

// -------------------------------------------------------------- fluxes --
// Face arrays have shape [5][nfz][nfy][nfx] with nf_d = n_d + (d == DIR).
// Lanes of a wave are consecutive in i, so every load and store is coalesced.
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
    // one face per thread: the face kernels are limited by their register
    // footprint (Riemann solver state), so they keep the shortest live ranges
    const int nfx = P.nx + (DIR == 1), nfy = P.ny + (DIR == 2), nfz = P.nz + (DIR == 3);
    const int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nfx * nfy * nfz) return;
    const int a = (int)(t % nfx);
    const int64_t rem = t / nfx;
    const int b = (int)(rem % nfy), c = (int)(rem / nfy);
    // reference indices: faces 0..n along DIR, cells 1..n across
    const int i = a + (DIR != 1), j = b + (DIR != 2), k = c + (DIR != 3);
    const int fi = (DIR == 1) ? i : (DIR == 2) ? j : k;
    const int nodes = (DIR == 1) ? P.nx : (DIR == 2) ? P.ny : P.nz;
    const double eps64 = 64.0 * DBL_EPSILON;
    bool zero = false;
    if (DIR == 1 && P.cyl) zero = fabs(xf(P, i)) <= eps64;           // evolution.f90:69-71
    if (DIR == 2 && P.polar) zero = fabs(sin(yf(P, j))) <= eps64;    // evolution.f90:122-124
    R out[NEQ];
    if (zero) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) out[v] = 0;
    } else {
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
        flux_face<MK, DIR, R>(P, p, f, i, j, k, fi, nodes, out);
    }
    const int64_t fo = ((int64_t)c * nfy + b) * nfx + a;
#pragma unroll
    for (int v = 0; v < NEQ; v++) F[v * fstride + fo] = out[v];
}

// --------------------------------------------------------- cell update --
// stage 1..3: rhs (evolution.f90:85-88,137-139,180-182) and the RK
// combination (grhd2.f90:243-249) into u; cons_to_prim follows in
// k_c2p_stage.  stage 0 writes the rhs itself to u (parity entry point).
// The metric depends on (r, theta) only; RUN > 1 lets a thread reuse its
// centre metric over RUN consecutive phi cells (fewer metric evaluations, but
// twice the registers: measured slower, so the launcher uses RUN = 1).
template <int MK, typename R, int RUN>
__global__ __launch_bounds__(256) void k_update(Params P, double *__restrict__ p,
                                                const R *__restrict__ Fx, const R *__restrict__ Fy,
                                                const R *__restrict__ Fz, int64_t fsx, int64_t fsy,
                                                int64_t fsz, const double *__restrict__ u0,
                                                double *__restrict__ u, int stage, double dt,
                                                unsigned long long *__restrict__ counters)
{
    const int nchunk = (P.nz + RUN - 1) / RUN;
    const int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= (int64_t)P.nx * P.ny * nchunk) return;
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    const int i0 = (int)(t % P.nx);
    const int64_t rem = t / P.nx;
    const int j0 = (int)(rem % P.ny), chunk = (int)(rem / P.ny);
    const int i = i0 + 1, j = j0 + 1;
    CentreMetric m;
    metric_centre<MK>(P, xc(P, i), yc(P, j), m);
    const int k0b = chunk * RUN, k0e = min(k0b + RUN, P.nz);
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
#pragma unroll
        for (int v = 0; v < NEQ; v++) u[v * ncell + tc] = c[v];
    }
}

// cons_to_prim after the RK combination (grhd2.f90:252-260), one cell per
// thread: a serial, divide-heavy root solve per cell, kept apart from the
// sources so that it runs at a higher occupancy.  WARM starts the root solve
// from the cell's previous Lorentz factor (mixed-precision path).
template <int MK, bool WARM>
__global__ __launch_bounds__(256) void k_c2p_stage(Params P, double *__restrict__ p,
                                                   double *__restrict__ u,
                                                   unsigned long long *__restrict__ counters)
{
    const int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    if (t >= ncell) return;
    const int i = (int)(t % P.nx) + 1;
    const int64_t rem = t / P.nx;
    const int j = (int)(rem % P.ny) + 1, k = (int)(rem / P.ny) + 1;
    double g[3][3], gi[3][3], sqg, alpha, beta[3], q[NEQ], c[NEQ];
    metric_c2p<MK>(P, xc(P, i), yc(P, j), g, gi, sqg, alpha, beta);
    const int64_t pc = pindex(P, i, j, k);
    double w_start = 0.0;
    if (WARM) {
        double v[3] = {p[2 * P.vstride + pc], p[3 * P.vstride + pc], p[4 * P.vstride + pc]};
        double vsq = 0.0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) vsq = vsq + g[a][b] * v[a] * v[b];
        w_start = (vsq < 1.0) ? 1.0 / sqrt(1.0 - vsq) : 0.0;
    }
#pragma unroll
    for (int v = 0; v < NEQ; v++) c[v] = u[v * ncell + t];
    int brent = 0;
    bool repaired = false;
    const bool atm = c2p(P.g1, P.rho_floor, P.p_floor, g, gi, sqg, c, q, brent, w_start, &repaired);
    if (atm) atomicAdd(&counters[0], 1ull);
    if (brent) atomicAdd(&counters[1], 1ull);
#pragma unroll
    for (int v = 0; v < NEQ; v++) p[v * P.vstride + pc] = q[v];
    // U changes only where con2prim rescales the momentum or resets the cell
    if (repaired) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) u[v * ncell + t] = c[v];
    }
}

template <int MK, typename R>
void launch_stage_mk(const Params &P, hipStream_t s, double *p, void *const F[3], const int64_t fstr[3],
                     const double *u0, double *u, int stage, double dt, unsigned long long *counters)
{
    const int bs = 256;
    const int64_t nfx = (int64_t)(P.nx + 1) * P.ny * P.nz;
    const int64_t nfy = (int64_t)P.nx * (P.ny + 1) * P.nz;
    const int64_t nfz = (int64_t)P.nx * P.ny * (P.nz + 1);
    constexpr int RUN = 1;  // one cell per thread: 4 waves/SIMD beat metric reuse along phi
    const int64_t ncol = (int64_t)P.nx * P.ny * ((P.nz + RUN - 1) / RUN);
    R *Fx = (R *)F[0], *Fy = (R *)F[1], *Fz = (R *)F[2];
    k_flux<MK, 1, R><<<blocks_for(nfx, bs), bs, 0, s>>>(P, p, Fx, fstr[0]);
    if (P.ny > 1) k_flux<MK, 2, R><<<blocks_for(nfy, bs), bs, 0, s>>>(P, p, Fy, fstr[1]);
    if (P.nz > 1) k_flux<MK, 3, R><<<blocks_for(nfz, bs), bs, 0, s>>>(P, p, Fz, fstr[2]);
    k_update<MK, R, RUN><<<blocks_for(ncol, bs), bs, 0, s>>>(P, p, Fx, Fy, Fz, fstr[0], fstr[1], fstr[2], u0,
                                                        u, stage, dt, counters);
    if (stage > 0) {
        const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
        constexpr bool warm = sizeof(R) == sizeof(float);
        k_c2p_stage<MK, warm><<<blocks_for(ncell, bs), bs, 0, s>>>(P, p, u, counters);
    }
}

// Instantiates the stage launcher of one face arithmetic for every metric kind.
#define GRHD_STAGE_DISPATCH(R)                                                                    \
    switch (P.mk) {                                                                               \
    case MK_MINK_CART: launch_stage_mk<MK_MINK_CART, R>(P, s, p, F, fstr, u0, u, stage, dt, counters); break; \
    case MK_MINK_CYL: launch_stage_mk<MK_MINK_CYL, R>(P, s, p, F, fstr, u0, u, stage, dt, counters); break;   \
    case MK_MINK_SPH: launch_stage_mk<MK_MINK_SPH, R>(P, s, p, F, fstr, u0, u, stage, dt, counters); break;   \
    case MK_MINK_SPH_LOG: launch_stage_mk<MK_MINK_SPH_LOG, R>(P, s, p, F, fstr, u0, u, stage, dt, counters); break; \
    case MK_EF: launch_stage_mk<MK_EF, R>(P, s, p, F, fstr, u0, u, stage, dt, counters); break;               \
    case MK_EF_LOG: launch_stage_mk<MK_EF_LOG, R>(P, s, p, F, fstr, u0, u, stage, dt, counters); break;       \
    case MK_KS: launch_stage_mk<MK_KS, R>(P, s, p, F, fstr, u0, u, stage, dt, counters); break;               \
    default: launch_stage_mk<MK_KS_LOG, R>(P, s, p, F, fstr, u0, u, stage, dt, counters); break;              \
    }

}  // namespace grhd
