// grhd_kernels.hip -- device kernels of the SSP-RK3 finite-volume update.
//
// One RK stage = one face kernel (every active direction) + one cell kernel
// (flux divergence, sources, RK combination, con2prim) + one boundary kernel.
// Kernels are templated on the metric kind so each launch evaluates its
// analytic metric with compile-time structure; reconstruction and the Riemann
// solver are uniform runtime branches.  This unit holds the boundary,
// conversion, CFL and observer kernels; the stage kernels are instantiated per
// metric kind by grhd_stage_{f64,f32}_*.hip.
#include <algorithm>

#include "grhd_stage.hpp"
#include "grhd_bound.hpp"

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
    metric_c2p<MK>(P, coord_cc(P, i, j), g, gi, sqg, alpha, beta);
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
    metric_c2p<MK>(P, coord_cc(P, i, j), g, gi, sqg, alpha, beta);
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


template <int MK>
__global__ __launch_bounds__(256) void k_boundaries(Params P, double *__restrict__ p,
                                                    const double *__restrict__ injector)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ghost_cells(P)) return;
    bound_body<MK>(P, p, injector, t);
}

// --------------------------------------------------------- mass monitor --
// Mass monitor (grhd2.f90:283-301): D dx dy dz summed over the interior, a
// grid-stride sum per thread, wave64 shuffle, then the block's waves through
// LDS; one partial per block, added on the host in block order.  D is the
// first plane of the SoA conserved array, so the reads are coalesced.
__global__ __launch_bounds__(256) void k_mass(Params P, const double *__restrict__ D, int64_t ncell,
                                              double *__restrict__ partial)
{
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ncell;
         t += (int64_t)gridDim.x * blockDim.x)
        acc = acc + D[t] * P.dx * P.dy * P.dz;
    for (int off = 32; off > 0; off >>= 1) acc = acc + __shfl_xor(acc, off, 64);
    __shared__ double red[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

int launch_mass(const Params &P, hipStream_t s, const double *D, double *partial, int max_blocks)
{
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(max_blocks, (ncell + 255) / 256));
    k_mass<<<nb, 256, 0, s>>>(P, D, ncell, partial);
    return nb;
}

// ------------------------------------------------------------ observers --
// Block reduction of one value by op (sum, min or max) over the 256 threads:
// wave64 shuffle, then the four waves through LDS.  Thread 0 gets the result.
enum { RED_SUM = 0, RED_MIN = 1, RED_MAX = 2 };
__device__ __forceinline__ double red_op(int op, double a, double b)
{
    return op == RED_SUM ? a + b : (op == RED_MIN ? fmin(a, b) : fmax(a, b));
}
__device__ double block_reduce(int op, double v, double *sh)
{
    for (int off = 32; off > 0; off >>= 1) v = red_op(op, v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();  // sh is reused by consecutive reductions
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    return red_op(op, red_op(op, sh[0], sh[1]), red_op(op, sh[2], sh[3]));
}

// calc_global_diagnostics (output.f90:1052-1125) over the interior: per block
// NGLOB partials -- mass, disk mass (rho > 1e3 rho_floor), angular momentum
// (S_3), rho min / max and p min / max over finite values, max v^2 over finite
// values, atmosphere and invalid cell counts.  v^2 = gamma_ij v^i v^j with the
// centre metric evaluated in registers.
constexpr int NGLOB = 10;
template <int MK>
__global__ __launch_bounds__(256) void k_global(Params P, const double *__restrict__ p,
                                                const double *__restrict__ u, double *__restrict__ partial)
{
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    const double vol = P.dx * P.dy * P.dz, cut = 1.0e3 * P.rho_floor;
    double acc[NGLOB] = {0.0, 0.0, 0.0, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, 0.0, 0.0, 0.0};
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ncell;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % P.nx) + 1;
        const int64_t rem = t / P.nx;
        const int j = (int)(rem % P.ny) + 1, k = (int)(rem / P.ny) + 1;
        const int64_t pc = pindex(P, i, j, k);
        double q[NEQ];
        for (int v = 0; v < NEQ; v++) q[v] = p[v * P.vstride + pc];
        double g[3][3], gi[3][3], sqg, alpha, beta[3];
        metric_c2p<MK>(P, coord_cc(P, i, j), g, gi, sqg, alpha, beta);
        double vsq = 0.0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) vsq = vsq + g[a][b] * q[EQ_VX + a] * q[EQ_VX + b];
        const double rho = q[EQ_DE], pr = q[EQ_PR];
        if (!isfinite(rho) || !isfinite(pr) || !isfinite(vsq) || rho <= 0.0 || pr <= 0.0 || vsq < 0.0 ||
            vsq >= 1.0)
            acc[9] = acc[9] + 1.0;
        if (isfinite(rho)) {
            acc[3] = fmin(acc[3], rho);
            acc[5] = fmax(acc[5], rho);
        }
        if (isfinite(pr)) {
            acc[4] = fmin(acc[4], pr);
            acc[6] = fmax(acc[6], pr);
        }
        if (isfinite(vsq)) acc[7] = fmax(acc[7], vsq);
        const double D = u[EQ_DE * ncell + t];
        acc[0] = acc[0] + D * vol;
        acc[2] = acc[2] + u[EQ_VZ * ncell + t] * vol;
        if (rho > cut) acc[1] = acc[1] + D * vol;
        else acc[8] = acc[8] + 1.0;
    }
    __shared__ double sh[4];
    const int ops[NGLOB] = {RED_SUM, RED_SUM, RED_SUM, RED_MIN, RED_MIN, RED_MAX, RED_MAX, RED_MAX, RED_SUM, RED_SUM};
    for (int q = 0; q < NGLOB; q++) {
        const double r = block_reduce(ops[q], acc[q], sh);
        if (threadIdx.x == 0) partial[(size_t)blockIdx.x * NGLOB + q] = r;
    }
}

// Mass of each phi plane above the density cut (calc_ppi_modes,
// output.f90:1011-1020): one block per plane k, partial[k - 1].
__global__ __launch_bounds__(256) void k_plane_mass(Params P, const double *__restrict__ p,
                                                    const double *__restrict__ u, double *__restrict__ partial)
{
    const int k = blockIdx.x + 1;
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz, plane = (int64_t)P.nx * P.ny;
    const double vol = P.dx * P.dy * P.dz, cut = 1.0e3 * P.rho_floor;
    double acc = 0.0;
    for (int64_t c = threadIdx.x; c < plane; c += blockDim.x) {
        const int i = (int)(c % P.nx) + 1, j = (int)(c / P.nx) + 1;
        if (p[pindex(P, i, j, k)] > cut) acc = acc + u[EQ_DE * ncell + (int64_t)(k - 1) * plane + c] * vol;
    }
    __shared__ double sh[4];
    const double r = block_reduce(RED_SUM, acc, sh);
    if (threadIdx.x == 0) partial[k - 1] = r;
}

// calc_m_dot (output.f90:951-993) on cell i = probe: D (alpha v^x - beta^x) dy dz
// summed over (j, k), one partial per block.
template <int MK>
__global__ __launch_bounds__(256) void k_mdot(Params P, int probe, const double *__restrict__ p,
                                              const double *__restrict__ u, double *__restrict__ partial)
{
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz, njk = (int64_t)P.ny * P.nz;
    double acc = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < njk; c += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(c % P.ny) + 1, k = (int)(c / P.ny) + 1;
        double g[3][3], gi[3][3], sqg, alpha, beta[3];
        metric_c2p<MK>(P, coord_cc(P, probe, j), g, gi, sqg, alpha, beta);
        const double veff = alpha * p[EQ_VX * P.vstride + pindex(P, probe, j, k)] - beta[0];
        const int64_t t = ((int64_t)(k - 1) * P.ny + (j - 1)) * P.nx + (probe - 1);
        acc = acc + u[EQ_DE * ncell + t] * veff * P.dy * P.dz;
    }
    __shared__ double sh[4];
    const double r = block_reduce(RED_SUM, acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
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
        metric_centre<MK>(P, coord_cc(P, i, j), m);
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
    k_boundaries<MK><<<blocks_for(ghost_cells(P), 256), 256, 0, s>>>(P, p, injector);
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
    // small and 2D grids: fused stage form (3 launches)
    const bool fused = single_stage_fused(P);
    if (mixed) launch_stage_f32(P, s, p, F, fstr, u0, u, stage, dt, counters, fused);
    else launch_stage_f64(P, s, p, F, fstr, u0, u, stage, dt, counters, fused);
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


int launch_global(const Params &P, hipStream_t s, const double *p, const double *u, double *partial, int max_blocks)
{
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(max_blocks, (ncell + 255) / 256));
    GRHD_DISPATCH(P.mk, (k_global<MKC><<<nb, 256, 0, s>>>(P, p, u, partial)));
    return nb;
}

void launch_plane_mass(const Params &P, hipStream_t s, const double *p, const double *u, double *partial)
{
    k_plane_mass<<<P.nz, 256, 0, s>>>(P, p, u, partial);
}

int launch_mdot(const Params &P, hipStream_t s, int probe, const double *p, const double *u, double *partial,
                int max_blocks)
{
    const int64_t njk = (int64_t)P.ny * P.nz;
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(max_blocks, (njk + 255) / 256));
    GRHD_DISPATCH(P.mk, (k_mdot<MKC><<<nb, 256, 0, s>>>(P, probe, p, u, partial)));
    return nb;
}

}  // namespace grhd
