// grhd_batch_kernels.hpp -- the RK3 stage of many independent problems in
// one launch.  Included by the two translation units that instantiate it
// (grhd_batch_f64.hip with FP64 faces, grhd_batch_mixed.hip with FP32 faces).
//
// A batch group holds problems of one metric kind and one face arithmetic, of
// any grid sizes.  Each phase of the stage (x, y, z faces; cells; ghost
// cells) is a list of logical blocks: problem r owns blocks
// [begin[r], begin[r+1]) and runs the same device body as the single-problem
// kernels on its own block index, so a problem gives bit-identical results
// alone or in any mix.  Two launch forms:
//   - kernels per phase, captured in a hipGraph: either faces of all
//     directions, cells (update and con2prim fused per thread) and ghosts
//     -- three launches per stage -- or one face launch per direction and
//     separate update and con2prim launches;
//   - one persistent kernel per stage: a co-resident grid walks the three
//     face phases, a grid barrier, the cell phase (update and con2prim fused
//     per thread), a barrier, and the ghost phase.
#pragma once
#include "grhd_batch.hpp"
#include "grhd_bound.hpp"
#include "grhd_stage.hpp"

namespace grhd {

// Problem of logical block L: the last r with begin[r] <= L (begin has n + 1
// entries, begin[n] = total).
__device__ __forceinline__ int find_run(const int64_t *begin, int n, int64_t L)
{
    int lo = 0, hi = n;  // begin[lo] <= L < begin[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (begin[mid] <= L) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <int DIR>
__device__ __forceinline__ int faces_z(const Params &P)
{
    return P.nz + (DIR == 3);
}

// --------------------------------------------------------- one phase each --
template <int MK, int DIR, typename R>
__device__ __forceinline__ void bflux_block(const BatchTables &T, const double *__restrict__ dts,
                                            int64_t L, R *lds)
{
    const int64_t *begin = T.begin + (DIR - 1) * (T.n + 1);
    const int r = find_run(begin, T.n, L);
    const int sl = T.slot[r];
    if (dts[sl] == 0.0) return;  // problem finished: uniform over the block
    const RunDesc &rd = T.runs[sl];
    const FluxTiling t = flux_tiling(faces_z<DIR>(rd.P), T.gw);
    flux_body<MK, DIR, R>(rd.P, rd.p, reinterpret_cast<R *>(rd.F[DIR - 1]), rd.fstr[DIR - 1],
                          L - begin[r], t.kt, t.gw, lds);
}

template <int MK, int DIR, typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void kb_flux(
    BatchTables T, const double *__restrict__ dts)
{
    extern __shared__ double lds_geo[];
    bflux_block<MK, DIR, R>(T, dts, xcd_block(blockIdx.x, gridDim.x), reinterpret_cast<R *>(lds_geo));
}

// Faces of every problem and direction: logical blocks are the x-face
// blocks of all problems, then the y and the z ones.
template <int MK, typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void kb_flux3(
    BatchTables T, const double *__restrict__ dts)
{
    extern __shared__ double lds_geo[];
    R *lds = reinterpret_cast<R *>(lds_geo);
    int64_t L = xcd_block(blockIdx.x, gridDim.x);
    if (L < T.total[PH_FX]) {
        bflux_block<MK, 1, R>(T, dts, L, lds);
        return;
    }
    L -= T.total[PH_FX];
    if (L < T.total[PH_FY]) {
        bflux_block<MK, 2, R>(T, dts, L, lds);
        return;
    }
    bflux_block<MK, 3, R>(T, dts, L - T.total[PH_FY], lds);
}

// Buffer holding U^n at step sstep of the chunk: runs alternate their two
// conserved buffers every step.
__device__ __forceinline__ double *u_base(const RunDesc &rd, int sstep)
{
    return rd.u[(rd.ucur + sstep) & 1];
}
__device__ __forceinline__ double *u_stage(const RunDesc &rd, int sstep)
{
    return rd.u[(rd.ucur + sstep + 1) & 1];
}

template <int MK, typename R>
__global__ __launch_bounds__(256) void kb_update(BatchTables T, const double *__restrict__ dts,
                                                 int sstep, int stage)
{
    const int64_t L = xcd_block(blockIdx.x, gridDim.x);
    const int64_t *begin = T.begin + PH_CELL * (T.n + 1);
    const int r = find_run(begin, T.n, L);
    const int sl = T.slot[r];
    const double dt = dts[sl];
    if (dt == 0.0) return;
    const RunDesc &rd = T.runs[sl];
    const Params &P = rd.P;
    const int64_t t = (L - begin[r]) * 256 + threadIdx.x;
    if (t >= (int64_t)P.nx * P.ny * P.nz) return;
    update_body<MK, R>(P, rd.p, reinterpret_cast<const R *>(rd.F[0]), reinterpret_cast<const R *>(rd.F[1]),
                       reinterpret_cast<const R *>(rd.F[2]), rd.fstr[0], rd.fstr[1], rd.fstr[2],
                       u_base(rd, sstep), u_stage(rd, sstep), stage, dt, t);
}

template <int MK, bool WARM>
__global__ __launch_bounds__(256) void kb_c2p(BatchTables T, const double *__restrict__ dts, int sstep)
{
    const int64_t L = xcd_block(blockIdx.x, gridDim.x);
    const int64_t *begin = T.begin + PH_CELL * (T.n + 1);
    const int r = find_run(begin, T.n, L);
    const int sl = T.slot[r];
    if (dts[sl] == 0.0) return;
    const RunDesc &rd = T.runs[sl];
    const int64_t t = (L - begin[r]) * 256 + threadIdx.x;
    if (t >= (int64_t)rd.P.nx * rd.P.ny * rd.P.nz) return;
    c2p_body<MK, WARM>(rd.P, rd.p, u_stage(rd, sstep), rd.counters, t);
}

// Cell update and con2prim of the same cell in one thread.
template <int MK, typename R>
__global__ __launch_bounds__(256) void kb_cell(BatchTables T, const double *__restrict__ dts, int sstep,
                                               int stage)
{
    const int64_t L = xcd_block(blockIdx.x, gridDim.x);
    const int64_t *begin = T.begin + PH_CELL * (T.n + 1);
    const int r = find_run(begin, T.n, L);
    const int sl = T.slot[r];
    const double dt = dts[sl];
    if (dt == 0.0) return;
    const RunDesc &rd = T.runs[sl];
    const Params &P = rd.P;
    const int64_t t = (L - begin[r]) * 256 + threadIdx.x;
    if (t >= (int64_t)P.nx * P.ny * P.nz) return;
    double *u1 = u_stage(rd, sstep);
    update_body<MK, R>(P, rd.p, reinterpret_cast<const R *>(rd.F[0]), reinterpret_cast<const R *>(rd.F[1]),
                       reinterpret_cast<const R *>(rd.F[2]), rd.fstr[0], rd.fstr[1], rd.fstr[2],
                       u_base(rd, sstep), u1, stage, dt, t);
    c2p_body<MK, sizeof(R) == sizeof(float)>(P, rd.p, u1, rd.counters, t);
}

template <int MK>
__global__ __launch_bounds__(256) void kb_bound(BatchTables T, const double *__restrict__ dts)
{
    const int64_t L = xcd_block(blockIdx.x, gridDim.x);
    const int64_t *begin = T.begin + PH_BOUND * (T.n + 1);
    const int r = find_run(begin, T.n, L);
    const int sl = T.slot[r];
    if (dts[sl] == 0.0) return;
    const RunDesc &rd = T.runs[sl];
    const int64_t t = (L - begin[r]) * 256 + threadIdx.x;
    if (t >= ghost_cells(rd.P)) return;
    bound_body<MK>(rd.P, rd.p, rd.injector, t);
}

// ------------------------------------------------------ persistent stage --
// Grid-wide barrier for a co-resident grid: arrive on bar[0], the last block
// resets it and advances the generation bar[1].  The wait is bounded by the
// 100 MHz constant clock: past the limit the block records the failure in
// *err and continues, so the kernel always ends and the host reports it.
// Once *err is set (by this or an earlier barrier of the captured chunk) no
// block waits any more, so the rest of the chunk drains at once.
__device__ __forceinline__ void grid_barrier(unsigned *bar, unsigned nblocks, unsigned *err,
                                             uint64_t limit_ticks)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned gen = __hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const unsigned arrived =
            __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == nblocks - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
            const uint64_t t0 = wall_clock64();
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(1);
                if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
                if (wall_clock64() - t0 > limit_ticks) {
                    atomicOr(err, 1u);
                    break;
                }
            }
        }
        __threadfence();
    }
    __syncthreads();
}

// Block b's contiguous share [lo, hi) of a phase with `total` logical blocks.
__device__ __forceinline__ void share(int64_t total, int64_t &lo, int64_t &hi)
{
    const int64_t per = (total + gridDim.x - 1) / gridDim.x;
    lo = (int64_t)blockIdx.x * per;
    hi = lo + per < total ? lo + per : total;
}

template <int MK, typename R>
__global__ __launch_bounds__(256) void kb_stage(BatchTables T, const double *__restrict__ dts, int sstep,
                                                int stage, unsigned *bar, unsigned *err,
                                                uint64_t limit_ticks)
{
    extern __shared__ double lds_geo[];
    R *lds = reinterpret_cast<R *>(lds_geo);
    int64_t lo, hi;
    share(T.total[PH_FX], lo, hi);
    for (int64_t L = lo; L < hi; L++) bflux_block<MK, 1, R>(T, dts, L, lds);
    share(T.total[PH_FY], lo, hi);
    for (int64_t L = lo; L < hi; L++) bflux_block<MK, 2, R>(T, dts, L, lds);
    share(T.total[PH_FZ], lo, hi);
    for (int64_t L = lo; L < hi; L++) bflux_block<MK, 3, R>(T, dts, L, lds);
    grid_barrier(bar, gridDim.x, err, limit_ticks);
    {
        const int64_t *begin = T.begin + PH_CELL * (T.n + 1);
        share(T.total[PH_CELL], lo, hi);
        for (int64_t L = lo; L < hi; L++) {
            const int r = find_run(begin, T.n, L);
            const int sl = T.slot[r];
            const double dt = dts[sl];
            if (dt == 0.0) continue;
            const RunDesc &rd = T.runs[sl];
            const Params &P = rd.P;
            const int64_t t = (L - begin[r]) * 256 + threadIdx.x;
            if (t >= (int64_t)P.nx * P.ny * P.nz) continue;
            double *u1 = u_stage(rd, sstep);
            update_body<MK, R>(P, rd.p, reinterpret_cast<const R *>(rd.F[0]),
                               reinterpret_cast<const R *>(rd.F[1]), reinterpret_cast<const R *>(rd.F[2]),
                               rd.fstr[0], rd.fstr[1], rd.fstr[2], u_base(rd, sstep), u1, stage, dt, t);
            c2p_body<MK, sizeof(R) == sizeof(float)>(P, rd.p, u1, rd.counters, t);
        }
    }
    grid_barrier(bar, gridDim.x, err, limit_ticks);
    {
        const int64_t *begin = T.begin + PH_BOUND * (T.n + 1);
        share(T.total[PH_BOUND], lo, hi);
        for (int64_t L = lo; L < hi; L++) {
            const int r = find_run(begin, T.n, L);
            const int sl = T.slot[r];
            if (dts[sl] == 0.0) continue;
            const RunDesc &rd = T.runs[sl];
            const int64_t t = (L - begin[r]) * 256 + threadIdx.x;
            if (t >= ghost_cells(rd.P)) continue;
            bound_body<MK>(rd.P, rd.p, rd.injector, t);
        }
    }
}

// ------------------------------------------------------- small problems --
// One block per problem advances it K whole steps in one launch: faces of
// every direction, update and con2prim, ghost fill, three stages a step, with
// block barriers between the phases.  For problems of a few thousand cells
// the step then costs the block's arithmetic instead of nine kernel launches,
// and many such problems run side by side on different CUs.  dts holds K rows
// of cap entries (row s: every slot's dt at step s, 0 once finished).
constexpr int kSmallThreads = 512;  // 8 waves: two per SIMD hide part of the latency

template <int MK, typename R>
__global__ __launch_bounds__(kSmallThreads) void kb_small(BatchTables T, const double *__restrict__ dts, int K,
                                                          int cap)
{
    extern __shared__ double lds_geo[];
    R *lds = reinterpret_cast<R *>(lds_geo);
    const int sl = T.slot[blockIdx.x];
    const RunDesc &rd = T.runs[sl];
    const Params &P = rd.P;
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz, nghost = ghost_cells(P);
    const FluxTiling tx = flux_tiling(faces_z<1>(P), T.gw), ty = flux_tiling(faces_z<2>(P), T.gw),
                     tz = flux_tiling(faces_z<3>(P), T.gw);
    constexpr int W = kSmallThreads / 64;
    const int64_t nbx = flux_blocks<1>(P, tx, W), nby = P.ny > 1 ? flux_blocks<2>(P, ty, W) : 0,
                  nbz = P.nz > 1 ? flux_blocks<3>(P, tz, W) : 0;
    R *Fx = reinterpret_cast<R *>(rd.F[0]), *Fy = reinterpret_cast<R *>(rd.F[1]),
      *Fz = reinterpret_cast<R *>(rd.F[2]);
    for (int s = 0; s < K; s++) {
        const double dt = dts[(size_t)s * cap + sl];
        if (dt == 0.0) break;  // finished: every later row is 0 as well
        double *u0 = u_base(rd, s), *u1 = u_stage(rd, s);
        for (int stage = 1; stage <= 3; stage++) {
            for (int64_t b = 0; b < nbx; b++) flux_body<MK, 1, R>(P, rd.p, Fx, rd.fstr[0], b, tx.kt, tx.gw, lds);
            for (int64_t b = 0; b < nby; b++) flux_body<MK, 2, R>(P, rd.p, Fy, rd.fstr[1], b, ty.kt, ty.gw, lds);
            for (int64_t b = 0; b < nbz; b++) flux_body<MK, 3, R>(P, rd.p, Fz, rd.fstr[2], b, tz.kt, tz.gw, lds);
            for (int64_t t = threadIdx.x; t < ncell; t += blockDim.x) {
                update_body<MK, R>(P, rd.p, Fx, Fy, Fz, rd.fstr[0], rd.fstr[1], rd.fstr[2], u0, u1, stage, dt, t);
                c2p_body<MK, sizeof(R) == sizeof(float)>(P, rd.p, u1, rd.counters, t);
            }
            __syncthreads();
            for (int64_t t = threadIdx.x; t < nghost; t += blockDim.x) bound_body<MK>(P, rd.p, rd.injector, t);
            __syncthreads();
        }
    }
}

// ----------------------------------------------------------------- launch --
template <int MK, typename R>
void batch_stage_mk(const BatchLaunch &B, hipStream_t s)
{
    const BatchTables &T = B.tables;
    const double *dts = B.dts;
    const size_t lds = (size_t)T.lds_groups * NFG * 64 * sizeof(R);
    if (B.small) {
        const size_t lds_small = (size_t)T.lds_groups * (kSmallThreads / 256) * NFG * 64 * sizeof(R);
        kb_small<MK, R><<<(unsigned)T.n, kSmallThreads, lds_small, s>>>(T, dts, B.K, B.cap);
        return;
    }
    if (B.persistent) {
        kb_stage<MK, R><<<B.grid, 256, lds, s>>>(T, dts, B.sstep, B.stage, B.bar, B.err, B.limit_ticks);
        return;
    }
    if (B.fused) {
        kb_flux3<MK, R><<<(unsigned)(T.total[PH_FX] + T.total[PH_FY] + T.total[PH_FZ]), 256, lds, s>>>(T, dts);
        kb_cell<MK, R><<<(unsigned)T.total[PH_CELL], 256, 0, s>>>(T, dts, B.sstep, B.stage);
    } else {
        if (T.total[PH_FX]) kb_flux<MK, 1, R><<<(unsigned)T.total[PH_FX], 256, lds, s>>>(T, dts);
        if (T.total[PH_FY]) kb_flux<MK, 2, R><<<(unsigned)T.total[PH_FY], 256, lds, s>>>(T, dts);
        if (T.total[PH_FZ]) kb_flux<MK, 3, R><<<(unsigned)T.total[PH_FZ], 256, lds, s>>>(T, dts);
        kb_update<MK, R><<<(unsigned)T.total[PH_CELL], 256, 0, s>>>(T, dts, B.sstep, B.stage);
        kb_c2p<MK, sizeof(R) == sizeof(float)><<<(unsigned)T.total[PH_CELL], 256, 0, s>>>(T, dts, B.sstep);
    }
    if (T.total[PH_BOUND]) kb_bound<MK><<<(unsigned)T.total[PH_BOUND], 256, 0, s>>>(T, dts);
}

template <int MK, typename R>
int batch_occupancy_mk(size_t lds)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kb_stage<MK, R>, 256, lds) != hipSuccess) return 0;
    return nb;
}

// Instantiates the batched stage of one face arithmetic R for the metric
// kinds of one half: PART a = Minkowski kinds (0-3), b = EF and KS (4-7).
#define GRHD_BATCH_CASE(MKV, R, CALL) case MKV: CALL(MKV, R); break;
#define GRHD_BATCH_STAGE_CALL(MKV, R) batch_stage_mk<MKV, R>(B, s)
#define GRHD_BATCH_OCC_CALL(MKV, R) return batch_occupancy_mk<MKV, R>(lds)
#define GRHD_BATCH_DISPATCH(R, PART, M0, M1, M2, M3)                                   \
    void batch_stage_##R##_##PART(const BatchLaunch &B, hipStream_t s)                \
    {                                                                                  \
        switch (B.mk) {                                                                \
            GRHD_BATCH_CASE(M0, R, GRHD_BATCH_STAGE_CALL)                              \
            GRHD_BATCH_CASE(M1, R, GRHD_BATCH_STAGE_CALL)                              \
            GRHD_BATCH_CASE(M2, R, GRHD_BATCH_STAGE_CALL)                              \
            GRHD_BATCH_CASE(M3, R, GRHD_BATCH_STAGE_CALL)                              \
        default: break;                                                                \
        }                                                                              \
    }                                                                                  \
    int batch_occupancy_##R##_##PART(int mk, size_t lds)                               \
    {                                                                                  \
        switch (mk) {                                                                  \
        case M0: return batch_occupancy_mk<M0, R>(lds);                                \
        case M1: return batch_occupancy_mk<M1, R>(lds);                                \
        case M2: return batch_occupancy_mk<M2, R>(lds);                                \
        case M3: return batch_occupancy_mk<M3, R>(lds);                                \
        default: return 0;                                                             \
        }                                                                              \
    }

}  // namespace grhd
