// grhd_stage.hpp -- templated kernels of one RK stage (face fluxes and the
// cell update) and their launcher.  Instantiated per metric kind by
// grhd_stage_f64_*.hip (FP64 faces, the reference's rounding sequence) and
// grhd_stage_f32_*.hip (FP32 faces).
#pragma once
#include <algorithm>
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

// Column (a, b) = (ab % nfx, ab / nfx) of face column ab = ab0 + lane, where
// ab0 is uniform over the wave: the division is done once per wave, in 32
// bits, and the lane offset wraps at most once when rows hold 64 or more
// faces (a per-lane 64-bit division otherwise).
__device__ __forceinline__ void face_column(int nfx, int64_t ab0, int lane, int &a, int &b)
{
    if (nfx >= 64 && ab0 + 64 < ((int64_t)1 << 31)) {
        const unsigned u0 = (unsigned)ab0;
        a = (int)(u0 % (unsigned)nfx) + lane;
        b = (int)(u0 / (unsigned)nfx);
        if (a >= nfx) {
            a -= nfx;
            b += 1;
        }
    } else {
        const int64_t ab = ab0 + lane;
        a = (int)(ab % nfx);
        b = (int)(ab / nfx);
    }
}

// Zero-based indices of an interior cell (x fastest).
struct CellIdx {
    int i0, j0, k0;
};
__device__ __forceinline__ CellIdx cell_of(const Params &P, int64_t t)
{
    const int64_t rem = t / P.nx;
    return {(int)(t % P.nx), (int)(rem % P.ny), (int)(rem / P.ny)};
}

// -------------------------------------------------------------- fluxes --
// Face arrays have shape [5][nfz][nfy][nfx] with nf_d = n_d + (d == DIR).

// The position-only part of face (i, j) into the LDS column col (field f at
// col[f * 64]): face metric in the face arithmetic (as calc_rhs passes it to
// the Riemann solver, evolution.f90:74-77, 110-114, 158-162), gamma^dd for
// the GRHD wave speeds (fluxes.f90:755-770), the HLLC frame and face speed
// (fluxes.f90:195-205, 402-473), and the zero-flux faces on the cylindrical
// axis and at the poles (evolution.f90:69-71, 122-124).
template <int MK, int DIR, typename R>
__device__ __forceinline__ void face_geometry(const Params &P, int i, int j, R *col, int64_t stride = 64)
{
    const double eps64 = 64.0 * DBL_EPSILON;
    bool zero = false;
    if (DIR == 1 && P.cyl) zero = fabs(xf(P, i)) <= eps64;
    if (DIR == 2 && P.polar) zero = fabs(tab_yf(P, j)[0]) <= eps64;  // |sin(theta_face)|
    int flags = zero ? FG_ZERO : 0;
    if (!zero) {
        FaceMetric m;
        if (DIR == 1) metric_face<MK>(P, coord_fc(P, i, j), m);
        else if (DIR == 2) metric_face<MK>(P, coord_cf(P, i, j), m);
        else metric_face<MK>(P, coord_cc(P, i, j), m);  // z faces: centre metric
        R g[3][3], beta[3];
        const R alpha = (R)m.alpha, sqg = (R)sqrtr(m.det);
        for (int x = 0; x < 3; x++) {
            beta[x] = (R)m.beta[x];
            for (int y = 0; y < 3; y++) g[x][y] = (R)m.g[x][y];
        }
        const R det = g[0][0] * (g[1][1] * g[2][2] - g[1][2] * g[2][1]) -
                      g[0][1] * (g[1][0] * g[2][2] - g[1][2] * g[2][0]) +
                      g[0][2] * (g[1][0] * g[2][1] - g[1][1] * g[2][0]);
        R gii;
        if (DIR == 1) gii = rdiv(g[1][1] * g[2][2] - g[1][2] * g[2][1], det);
        else if (DIR == 2) gii = rdiv(g[0][0] * g[2][2] - g[0][2] * g[2][0], det);
        else gii = rdiv(g[0][0] * g[1][1] - g[0][1] * g[1][0], det);
        R cof[3][3], gdd = 0;
        const bool ok = !(alpha <= (R)0 || sqg <= (R)0 || !finite(alpha) || !finite(sqg)) &&
                        face_frame<R, DIR>(g, cof, gdd);
        const R w = ok ? rdiv(beta[DIR - 1], alpha * sqrtr(gdd)) : (R)0;
        if (ok) flags |= FG_FRAME_OK;
        col[FG_FAC * stride] = sqg * alpha * sqrtr(gdd);
        col[FG_ALPHA * stride] = alpha;
        for (int x = 0; x < 3; x++) col[(FG_BETA + x) * stride] = beta[x];
        for (int x = 0; x < 3; x++)
            for (int y = x; y < 3; y++) col[(FG_G + sym_index(x, y)) * stride] = g[x][y];
        col[FG_SQG * stride] = sqg;
        for (int x = 0; x < 3; x++)
            for (int y = 0; y < 3; y++) col[(FG_COF + 3 * x + y) * stride] = ok ? cof[x][y] : (R)0;
        col[FG_GDD * stride] = gdd;
        col[FG_W * stride] = w;
        col[FG_GII * stride] = gii;
    }
    col[FG_FLAGS * stride] = (R)flags;
}

// The stencil of face fi along DIR (cells fi-2 .. fi+3 of row (i, j, k)):
// s[v][o + 2] = primitive v at offset o from the face's left cell.
template <int MK, int DIR, typename R, int SCH = 0>
__device__ __forceinline__ void load_stencil(const Params &P, const double *__restrict__ p, int i, int j, int k,
                                             int fi, R s[NEQ][6])
{
    const int ng = SCH ? ghost_depth(kSchemes[SCH].rec) : P.ng;
    const int lo = (ng >= 2) ? -2 : 0, hi = (ng >= 3) ? 3 : (ng == 2 ? 2 : 1);
    const int64_t step = (DIR == 1) ? 1 : (DIR == 2) ? (int64_t)P.NXs : (int64_t)P.NXs * P.NYs;
    const int64_t base = pindex(P, DIR == 1 ? fi : i, DIR == 2 ? fi : j, DIR == 3 ? fi : k);
    // Offsets outside the ghost depth read a clamped (valid) address and are
    // then zeroed, so all 30 loads issue together without branches.
#pragma unroll
    for (int o = -2; o <= 3; o++) {
        const bool in = (o >= lo && o <= hi);
        const int64_t off = base + (int64_t)(o < lo ? lo : (o > hi ? hi : o)) * step;
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
            const R x = (R)p[v * P.vstride + off];
            s[v][o + 2] = in ? x : (R)0;
        }
    }
}

// Face states q_L, q_R of face fi along DIR: stencil loads and
// reconstruct_1d_core.
template <int MK, int DIR, typename R, int SCH = 0>
__device__ __forceinline__ void face_states(const Params &P, const double *__restrict__ p, int i, int j, int k,
                                            int fi, int nodes, R qL[NEQ], R qR[NEQ])
{
    R s[NEQ][6];
    load_stencil<MK, DIR, R, SCH>(P, p, i, j, k, fi, s);
    recon_face<R, SCH>(P, s, fi, nodes, DIR, qL, qR);
}

// Flux of one face (reconstruction + resolve_riemann_problem, fluxes.f90:56-72).
// FP64 faces of a compiled scheme with HLLC: the HLLE fallback (19 of 1.6
// million faces on the 3D torus) recomputes the face states instead of
// keeping them live through the HLLC solve, which frees their registers there
// (those kernels then spill nothing at 128 VGPRs; +7 % on the 3D torus); the
// pointer is made opaque so that the recomputation is not merged with the
// first one.  Both are the same pure function of p: the fallback sees the
// same states.  The FP32 face kernels measured slower with it, and the
// run-time scheme (every reconstruction inlined twice) spills more with it
// (48 -> 224 B per lane): both keep the states.
template <int MK, int DIR, typename R, int SCH = 0, typename FG>
__device__ __forceinline__ void flux_face(const Params &P, const double *__restrict__ p,
                                          const FG &f, int i, int j, int k, int fi, int nodes,
                                          R out[NEQ])
{
    R qL[NEQ], qR[NEQ];
    face_states<MK, DIR, R, SCH>(P, p, i, j, k, fi, nodes, qL, qR);
    const int rs = SCH ? kSchemes[SCH].rs : P.rs;
    if constexpr (sizeof(R) != sizeof(double) || SCH == 0) {
        riemann<R, DIR, kSRHD<MK>>(rs, (R)P.g1, (R)P.gam, qL, qR, f, out);
        return;
    }
    if (rs == 2) {
        if (hllc_try<R, DIR>((R)P.g1, (R)P.gam, qL, qR, f, out)) return;
        const double *pr = p;
        asm volatile("" : "+v"(pr));
        face_states<MK, DIR, R, SCH>(P, pr, i, j, k, fi, nodes, qL, qR);
    }
    hlle<R, DIR, kSRHD<MK>>((R)P.g1, (R)P.gam, qL, qR, f, out);
}

// The faces of the compiled MP5 + sensor + HLLC scheme (SCH 1) whose HLLC
// solve succeeds (all but 19 of the 1.6 million faces of the 3D torus).
// Returns false, with out unset, where HLLC falls back to HLLE: those faces
// are computed by flux_face in a second kernel over a list of them
// (k_flux_defer), so the face kernel carries neither the fallback's second
// reconstruction and HLLE code nor the registers they need (FP64: no spill at
// 128 VGPRs, against 52 B per lane with the fallback inline; the z faces of
// the 3D torus 11 % and the x faces 5 % faster).  A face computed here gets
// flux_face's result bit for bit: the same stencil, reconstruction and
// solver.  Deferring the sensor's TVD-MC and WENO3 faces too (8-14 % of the
// x and y faces of the 3D torus) measured slower: the list is then long and
// its one counter contended.
template <int MK, int DIR, typename R, typename FG>
__device__ __forceinline__ bool flux_face_hot(const Params &P, const double *__restrict__ p, const FG &f, int i,
                                              int j, int k, int fi, int nodes, R out[NEQ])
{
    static_assert(kSchemes[1].rec == 4 && kSchemes[1].sensor && kSchemes[1].rs == 2, "scheme 1");
    R qL[NEQ], qR[NEQ];
    face_states<MK, DIR, R, 1>(P, p, i, j, k, fi, nodes, qL, qR);
    return hllc_try<R, DIR>((R)P.g1, (R)P.gam, qL, qR, f, out);
}

// Appends the calling lanes' faces with defer set to list (one atomic per
// wave: the wave's count is added once, and each lane writes at the counter's
// old value plus the number of deferring lanes below it).  Every active lane
// of the wave must call it.
__device__ __forceinline__ void defer_push(unsigned *__restrict__ ctr, unsigned *__restrict__ list, unsigned cap,
                                           unsigned face, bool defer)
{
    const unsigned long long m = __ballot(defer);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(ctr, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    if (defer) {
        const unsigned pos = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
        if (pos < cap) list[pos] = face;
    }
}

// Tiling of the face kernels: a tile is 64 consecutive faces of one (x, y)
// plane (flattened, so rows need not be multiples of 64) times kt planes along
// z.  The face geometry depends on (x, y) only, so a group of gw = kt waves
// shares one LDS copy of it: the group's first wave evaluates it and each wave
// of the group takes one plane.  A 256-thread block holds 4 / gw groups; grids
// with fewer than 4 planes use groups of one wave.
struct FluxTiling {
    int kt, gw;
};
__host__ __device__ inline FluxTiling flux_tiling(int nfz, int gw_pref = 4)
{
    if (nfz >= gw_pref) return {gw_pref, gw_pref};
    return {1, 1};
}

// Faces of logical block blk of one problem (0 <= blk < flux_blocks()).
// lds holds (waves / gw) * NFG * 64 values of R, waves = blockDim.x / 64.
template <int MK, int DIR, typename R, int SCH = 0>
__device__ __forceinline__ void flux_body(const Params &P, const double *__restrict__ p,
                                          R *__restrict__ F, int64_t fstride, int64_t blk, int kt,
                                          int gw, R *lds)
{
    const int nfx = P.nx + (DIR == 1), nfy = P.ny + (DIR == 2), nfz = P.nz + (DIR == 3);
    const int64_t nab = (int64_t)nfx * nfy;
    const int64_t nchunk = (nab + 63) / 64;
    const int groups = (int)(blockDim.x >> 6) / gw;
    const int64_t ncg = (nchunk + groups - 1) / groups;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int grp = wave / gw, wig = wave % gw;
    // Block order: the stencil of x and y faces reaches neighbouring rows of
    // the same plane, so consecutive blocks walk the plane (tiles of 64 faces)
    // first; the z stencil reaches neighbouring planes, so z-face blocks walk
    // along z first.  Consecutive blocks share an XCD's L2 (xcd_block).
    const int nct = (nfz + kt - 1) / kt;
    const int64_t cg = (DIR == 3) ? blk / nct : blk % ncg;
    const int ctile = (DIR == 3) ? (int)(blk % nct) : (int)(blk / ncg);
    const int64_t chunk = cg * groups + grp;
    const int64_t ab = chunk * 64 + lane;
    const bool active = ab < nab;
    int a = 0, b = 0;
    if (active) face_column(nfx, chunk * 64, lane, a, b);
    // reference indices: faces 0..n along DIR, cells 1..n across
    const int i = a + (DIR != 1), j = b + (DIR != 2);
    R *col = lds + grp * NFG * 64 + lane;
    if (wig == 0 && active) face_geometry<MK, DIR, R>(P, i, j, col);
    __syncthreads();
    // one plane per wave: a loop over planes would keep the geometry loads
    // hoisted in registers (more VGPRs, fewer waves)
    const int c = ctile * kt + wig;
    if (active && c < nfz) {
        const FaceGeo<R> fg{col};
        const int nodes = (DIR == 1) ? P.nx : (DIR == 2) ? P.ny : P.nz;
        const int k = c + (DIR != 3);
        const int fi = (DIR == 1) ? i : (DIR == 2) ? j : k;
        R out[NEQ];
        if (fg.flags() & FG_ZERO) {
#pragma unroll
            for (int v = 0; v < NEQ; v++) out[v] = 0;
        } else {
            flux_face<MK, DIR, R, SCH>(P, p, fg, i, j, k, fi, nodes, out);
        }
        const int64_t fo = (int64_t)c * nab + ab;
#pragma unroll
        for (int v = 0; v < NEQ; v++) F[v * fstride + fo] = out[v];
    }
    __syncthreads();  // the LDS columns are reused by the block's next tile
}

// Logical blocks of the faces of direction DIR for blocks of `waves` waves.
template <int DIR>
__host__ __device__ inline int64_t flux_blocks(const Params &P, FluxTiling t, int waves = 4)
{
    const int nfx = P.nx + (DIR == 1), nfy = P.ny + (DIR == 2), nfz = P.nz + (DIR == 3);
    const int groups = waves / t.gw;
    const int64_t nchunk = ((int64_t)nfx * nfy + 63) / 64;
    return ((nchunk + groups - 1) / groups) * ((nfz + t.kt - 1) / t.kt);
}

// Four waves per SIMD: the compiler keeps the face kernels at 128 VGPRs with a
// few doubles of scratch spill, measured faster than 151 VGPRs at three waves.
template <int MK, int DIR, typename R, int SCH>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_flux(
    Params P, const double *__restrict__ p, R *__restrict__ F, int64_t fstride, int kt, int gw)
{
    extern __shared__ double lds_geo[];
    flux_body<MK, DIR, R, SCH>(P, p, F, fstride, xcd_block(blockIdx.x, gridDim.x), kt, gw,
                               reinterpret_cast<R *>(lds_geo));
}

// Face geometry of every face column of direction DIR into the table tab
// (field f of column ab at tab[f * nab + ab]), once per stage.
template <int MK, int DIR, typename R>
__global__ __launch_bounds__(256) void k_geo(Params P, R *__restrict__ tab)
{
    const int nfx = P.nx + (DIR == 1), nfy = P.ny + (DIR == 2);
    const int64_t nab = (int64_t)nfx * nfy;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nab) return;
    const int a = (int)(t % nfx), b = (int)(t / nfx);
    face_geometry<MK, DIR, R>(P, a + (DIR != 1), b + (DIR != 2), tab + t, nab);
}

// Faces of direction DIR reading their geometry from the stage's table: the
// blocks, lanes and planes of flux_body with groups of four, with neither the
// geometry evaluation nor the barrier.  The z faces walk their blocks along z
// first (their stencil reaches the neighbouring planes).  x and y faces walk
// the plane first, or with kFaceOrderZ in their face form z first: each XCD
// then keeps the table rows of its range of face columns in L2 (the x and y
// tables of the 3D torus are 6.6 MB each, above an XCD's 4 MB), at the cost
// of the p rows the y stencil shares between neighbouring column tiles.  With
// kFaceOrderBand they walk the plane tiles of a band, then the next column
// tile, band after band, the bands being an eighth of the plane tiles: each
// XCD's contiguous range of blocks (xcd_block) is then about one band, a slab
// of planes whose p rows and table rows it reads once into its own L2.  On
// the 3D torus (bands of 4 plane tiles, 16 planes per XCD) the x and y faces
// measured 16 and 11 % faster than plane first; bands of 2, 3, 8 and 16 plane
// tiles were 1.4-5 % slower per stage than bands of 4.  DEFER (scheme 1 only): faces whose
// HLLC solve succeeds here (flux_face_hot), the HLLC fallbacks appended to
// the direction's list for k_flux_defer.
template <int MK, int DIR, typename R, int SCH, bool DEFER = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_flux_tab(
    Params P, const double *__restrict__ p, R *__restrict__ F, int64_t fstride, const R *__restrict__ tab)
{
    static_assert(!DEFER || SCH == 1, "deferred faces: scheme 1 only");
    const int nfx = P.nx + (DIR == 1), nfy = P.ny + (DIR == 2), nfz = P.nz + (DIR == 3);
    const int64_t nab = (int64_t)nfx * nfy;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // block -> (column tile, plane tile): uniform, in 32 bits (the grid has
    // fewer than 2^32 blocks), which keeps the scalar prologue short
    const unsigned nchunk = (unsigned)((nab + 63) / 64);
    const unsigned nct = (unsigned)(nfz + 3) / 4u;
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    const int order = (DIR == 3) ? kFaceOrderZ : (P.face_gw[DIR - 1] & kFaceOrders);
    unsigned cg, ctile;
    if (order == kFaceOrderZ) {
        cg = blk / nct;
        ctile = blk % nct;
    } else if (order == kFaceOrderBand) {
        const unsigned bw = (nct + kXCD - 1) / kXCD;  // plane tiles per band
        const unsigned band = blk / (nchunk * bw);
        const unsigned r = blk - band * nchunk * bw;
        const unsigned nb = min(bw, nct - band * bw);  // plane tiles of this band
        cg = r / nb;
        ctile = band * bw + r % nb;
    } else {
        cg = blk % nchunk;
        ctile = blk / nchunk;
    }
    const int64_t ab = (int64_t)cg * 64 + lane;
    const int c = (int)ctile * 4 + wave;
    if (ab >= nab || c >= nfz) return;
    int a, b;
    face_column(nfx, (int64_t)cg * 64, lane, a, b);
    const int i = a + (DIR != 1), j = b + (DIR != 2);
    const FaceGeo<R> fg{tab + ab, nab};
    const int nodes = (DIR == 1) ? P.nx : (DIR == 2) ? P.ny : P.nz;
    const int k = c + (DIR != 3);
    const int fi = (DIR == 1) ? i : (DIR == 2) ? j : k;
    const int64_t fo = (int64_t)c * nab + ab;
    R out[NEQ];
    bool done = true;
    if (fg.flags() & FG_ZERO) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) out[v] = 0;
    } else if constexpr (DEFER) {
        done = flux_face_hot<MK, DIR>(P, p, fg, i, j, k, fi, nodes, out);
    } else {
        flux_face<MK, DIR, R, SCH>(P, p, fg, i, j, k, fi, nodes, out);
    }
    if constexpr (DEFER) {
        defer_push(P.defer_ctr + 2 * (DIR - 1), P.defer_list[DIR - 1], (unsigned)(nab * nfz), (unsigned)fo, !done);
        if (!done) return;
    }
#pragma unroll
    for (int v = 0; v < NEQ; v++) F[v * fstride + fo] = out[v];
}

// The deferred faces of one direction (k_flux_tab with DEFER): each thread
// takes list entries q = thread, thread + grid, ... below the count and
// computes the face's HLLE flux from its reconstructed states -- what
// flux_face computes for a face whose HLLC solve fails (fluxes.f90:56-72),
// without repeating that solve.  Every block reads the same count (written by
// the face kernel before this launch) and so the same number of working
// blocks, enough for one entry per thread up to the grid: the others return
// at once, as every block does for an empty list.  The working block that
// finishes last clears the count and the block counter for the next stage:
// every working block has read the count before it adds itself to the block
// counter.
template <int MK, int DIR, typename R, int SCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_flux_defer(Params P, const double *__restrict__ p, R *__restrict__ F,
                                                    int64_t fstride, const R *__restrict__ tab)
{
    unsigned *ctr = P.defer_ctr + 2 * (DIR - 1);
    const unsigned *list = P.defer_list[DIR - 1];
    const int nfx = P.nx + (DIR == 1), nfy = P.ny + (DIR == 2);
    const int64_t nab = (int64_t)nfx * nfy;
    const int nodes = (DIR == 1) ? P.nx : (DIR == 2) ? P.ny : P.nz;
    // entries stop at the list's capacity (the face kernel writes none beyond)
    const unsigned cap = (unsigned)(nab * (P.nz + (DIR == 3)));
    const unsigned n = min(__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), cap);
    const unsigned nbk = min(gridDim.x, (n + blockDim.x - 1) / blockDim.x);  // working blocks
    if (blockIdx.x >= nbk) return;
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += nbk * blockDim.x) {
        // face indices are below 2^32 where the lists exist
        const unsigned fo = list[q];
        const int c = (int)(fo / (unsigned)nab);
        const unsigned ab = fo - (unsigned)c * (unsigned)nab;
        const int a = (int)(ab % (unsigned)nfx), b = (int)(ab / (unsigned)nfx);
        const int i = a + (DIR != 1), j = b + (DIR != 2), k = c + (DIR != 3);
        const int fi = (DIR == 1) ? i : (DIR == 2) ? j : k;
        const FaceGeo<R> fg{tab + ab, nab};
        R qL[NEQ], qR[NEQ], out[NEQ];
        face_states<MK, DIR, R, SCH>(P, p, i, j, k, fi, nodes, qL, qR);
        hlle<R, DIR, kSRHD<MK>>((R)P.g1, (R)P.gam, qL, qR, fg, out);
#pragma unroll
        for (int v = 0; v < NEQ; v++) F[v * fstride + fo] = out[v];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned fin = __hip_atomic_fetch_add(ctr + 1, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (fin == nbk - 1) {
            __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(ctr + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Blocks of k_flux_defer and k_xcell_fix: one pass over lists of up to
// 262,144 entries (the HLLC fallbacks of the 3D torus in FP64: none of the x
// faces, 1-3 % of the y and z faces, 40-140 thousand per stage), four waves
// per SIMD; longer lists take more passes.  The blocks past a list's extent
// return at once.
constexpr unsigned kDeferBlocks = 1024;

template <int MK, int DIR, typename R, int SCH>
void launch_flux(const Params &P, hipStream_t s, const double *p, R *F, int64_t fstr)
{
    if (face_form_table(P.face_gw[DIR - 1]) && P.geo_tab[DIR - 1]) {
        R *tab = static_cast<R *>(P.geo_tab[DIR - 1]);
        const int nfx = P.nx + (DIR == 1), nfy = P.ny + (DIR == 2), nfz = P.nz + (DIR == 3);
        const int64_t nab = (int64_t)nfx * nfy;
        if (!P.geo_ready) k_geo<MK, DIR, R><<<blocks_for(nab, 256), 256, 0, s>>>(P, tab);
        const int64_t nb = ((nab + 63) / 64) * ((nfz + 3) / 4);
        if constexpr (SCH == 1) {
            if (face_form_defers(P.face_gw[DIR - 1]) && P.defer_list[DIR - 1]) {
                k_flux_tab<MK, DIR, R, SCH, true><<<(unsigned)nb, 256, 0, s>>>(P, p, F, fstr, tab);
                k_flux_defer<MK, DIR, R, SCH><<<kDeferBlocks, 256, 0, s>>>(P, p, F, fstr, tab);
                return;
            }
        }
        k_flux_tab<MK, DIR, R, SCH><<<(unsigned)nb, 256, 0, s>>>(P, p, F, fstr, tab);
        return;
    }
    const int nfz = P.nz + (DIR == 3);
    const FluxTiling t = flux_tiling(nfz, (P.face_gw[DIR - 1] >= 1 && P.face_gw[DIR - 1] <= 8) ? P.face_gw[DIR - 1] : 4);
    const int waves = t.gw > 4 ? t.gw : 4;
    const size_t lds = (size_t)(waves / t.gw) * NFG * 64 * sizeof(R);
    k_flux<MK, DIR, R, SCH><<<(unsigned)flux_blocks<DIR>(P, t, waves), 64 * waves, lds, s>>>(P, p, F, fstr, t.kt,
                                                                                             t.gw);
}

// The compiled face scheme for a problem: the Kerr-Schild torus
// configurations get their own kernels, everything else the run-time one.
template <int MK>
inline int face_scheme(const Params &P)
{
    if constexpr (MK == MK_KS || MK == MK_KS_LOG) {
        for (int sch = 1; sch < (int)(sizeof(kSchemes) / sizeof(kSchemes[0])); sch++)
            if (P.rec == kSchemes[sch].rec && (P.sensor != 0) == (kSchemes[sch].sensor != 0) &&
                P.rs == kSchemes[sch].rs)
                return sch;
    }
    return 0;
}

template <int MK, typename R, int SCH>
void launch_faces(const Params &P, hipStream_t s, const double *p, R *Fx, R *Fy, R *Fz, const int64_t fstr[3])
{
    launch_flux<MK, 1, R, SCH>(P, s, p, Fx, fstr[0]);
    if (P.ny > 1) launch_flux<MK, 2, R, SCH>(P, s, p, Fy, fstr[1]);
    if (P.nz > 1) launch_flux<MK, 3, R, SCH>(P, s, p, Fz, fstr[2]);
}

// The face kernel of one direction (1..3) alone, with the problem's compiled
// face scheme (the face-tiling tuner times it; it writes only the face array).
template <int MK, typename R>
void launch_face_dir_mk(const Params &P, hipStream_t s, const double *p, void *const F[3], const int64_t fstr[3],
                        int dir)
{
    R *Fd = (R *)F[dir - 1];
    const int sch = face_scheme<MK>(P);
#define GRHD_FACE_DIR(SCHV)                                                        \
    do {                                                                           \
        if (dir == 1) launch_flux<MK, 1, R, SCHV>(P, s, p, Fd, fstr[0]);           \
        else if (dir == 2) launch_flux<MK, 2, R, SCHV>(P, s, p, Fd, fstr[1]);      \
        else launch_flux<MK, 3, R, SCHV>(P, s, p, Fd, fstr[2]);                    \
    } while (0)
    if constexpr (MK == MK_KS || MK == MK_KS_LOG) {
        if (sch == 1) { GRHD_FACE_DIR(1); return; }
        if (sch == 2) { GRHD_FACE_DIR(2); return; }
    }
    GRHD_FACE_DIR(0);
#undef GRHD_FACE_DIR
}

// --------------------------------------------------------- cell update --
// Cell t (0 <= t < nx ny nz, x fastest) of one problem.  stage 1..3: rhs
// (evolution.f90:85-88, 137-139, 180-182) and the RK combination
// (grhd2.f90:243-249) into u; cons_to_prim follows in c2p_body.  stage 0
// writes the rhs itself to u (parity entry point).  ZREG: the z-face fluxes
// of the cell are zl (left) and zr (right), held in registers by the caller,
// instead of entries of Fz.  XREG: the same for the x-face fluxes (xl, xr)
// on 3D grids; Fx is not read.
template <int MK, typename R, bool ZREG = false, bool XREG = false>
__device__ __forceinline__ void update_body(const Params &P, const double *__restrict__ p,
                                            const R *__restrict__ Fx, const R *__restrict__ Fy,
                                            const R *__restrict__ Fz, int64_t fsx, int64_t fsy,
                                            int64_t fsz, const double *__restrict__ u0,
                                            double *__restrict__ u, int stage, double dt, int64_t t,
                                            const R *zl = nullptr, const R *zr = nullptr,
                                            const R *xl = nullptr, const R *xr = nullptr,
                                            const CellIdx *ci = nullptr)
{
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    const CellIdx c0 = ci ? *ci : cell_of(P, t);
    const int i0 = c0.i0, j0 = c0.j0, k0 = c0.k0;
    const int i = i0 + 1, j = j0 + 1, k = k0 + 1;
    const int64_t pc = pindex(P, i, j, k);
    double q[NEQ], rhs[NEQ], src[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) q[v] = p[v * P.vstride + pc];
    if (MK != MK_MINK_CART) {
        CentreMetric m;
        metric_centre<MK>(P, coord_cc(P, i, j), m);
        sources(P.g1, m, q, src);
    } else {
#pragma unroll
        for (int v = 0; v < NEQ; v++) src[v] = 0.0;
    }
    // Every face value and both RK operands are loaded in one batch, so their
    // latencies overlap: the addresses of an inactive direction fall back to
    // the x faces (read and discarded), which keeps the loads branch-free.
    const bool ya = P.ny > 1, za = P.nz > 1;
    const int64_t fax = ((int64_t)k0 * P.ny + j0) * (P.nx + 1) + i0;
    const int64_t fay = ya ? ((int64_t)k0 * (P.ny + 1) + j0) * P.nx + i0 : fax;
    const R *Fyp = ya ? Fy : Fx, *Fzp = za ? Fz : Fx;
    const int64_t fsyp = ya ? fsy : fsx, fszp = za ? fsz : fsx;
    const int64_t sy = ya ? (int64_t)P.nx : 1, faz = za ? t : fax, sz = za ? (int64_t)P.nx * P.ny : 1;
    R fx0[NEQ], fx1[NEQ], fy0[NEQ], fy1[NEQ], fz0[NEQ], fz1[NEQ];
    double a0[NEQ], a1[NEQ];
    // stage 1 does not read u (its a1 loads repeat the u0 addresses and hit
    // the cache); stage 0 reads neither operand
    const bool rk = stage > 0;
    const int64_t tr = rk ? t : 0;
    const double *u0p = rk ? u0 : u, *u1p = (stage >= 2) ? u : u0p;
#pragma unroll
    for (int v = 0; v < NEQ; v++) {
        if constexpr (XREG) {
            fx0[v] = xl[v];
            fx1[v] = xr[v];
        } else {
            fx0[v] = Fx[v * fsx + fax];
            fx1[v] = Fx[v * fsx + fax + 1];
        }
        fy0[v] = Fyp[v * fsyp + fay];
        fy1[v] = Fyp[v * fsyp + fay + sy];
        if constexpr (ZREG) {
            fz0[v] = zl[v];
            fz1[v] = zr[v];
        } else {
            fz0[v] = Fzp[v * fszp + faz];
            fz1[v] = Fzp[v * fszp + faz + sz];
        }
        a0[v] = u0p[v * ncell + tr];
        a1[v] = u1p[v * ncell + tr];
    }
#pragma unroll
    for (int v = 0; v < NEQ; v++) rhs[v] = -quot((double)fx1[v] - (double)fx0[v], P.dx, P.inv_dx) + src[v];
    if (ya) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) rhs[v] = rhs[v] - quot((double)fy1[v] - (double)fy0[v], P.dy, P.inv_dy);
    }
    if (za) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) rhs[v] = rhs[v] - quot((double)fz1[v] - (double)fz0[v], P.dz, P.inv_dz);
    }
    if (stage == 0) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) u[v * ncell + t] = rhs[v];
        return;
    }
    double c[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) {
        if (stage == 1) c[v] = a0[v] + rhs[v] * dt;
        else if (stage == 2) c[v] = 0.75 * a0[v] + 0.25 * (a1[v] + rhs[v] * dt);
        else c[v] = quot(a0[v] + 2.0 * (a1[v] + rhs[v] * dt), 3.0, 1.0 / 3.0);
    }
#pragma unroll
    for (int v = 0; v < NEQ; v++) u[v * ncell + t] = c[v];
}

template <int MK, typename R>
__global__ __launch_bounds__(256) void k_update(Params P, const double *__restrict__ p,
                                                const R *__restrict__ Fx, const R *__restrict__ Fy,
                                                const R *__restrict__ Fz, int64_t fsx, int64_t fsy,
                                                int64_t fsz, const double *__restrict__ u0,
                                                double *__restrict__ u, int stage, double dt)
{
    const int64_t t = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= (int64_t)P.nx * P.ny * P.nz) return;
    update_body<MK, R>(P, p, Fx, Fy, Fz, fsx, fsy, fsz, u0, u, stage, dt, t);
}

// cons_to_prim after the RK combination (grhd2.f90:252-260) of cell t: a
// serial, divide-heavy root solve per cell.  WARM starts the root solve from
// the cell's previous Lorentz factor (mixed-precision path).  The primitives
// are written to p in place, or to pout with SPLIT (p is then only read).
template <int MK, bool WARM, bool SPLIT = false>
__device__ __forceinline__ void c2p_body(const Params &P, double *__restrict__ p,
                                         double *__restrict__ u,
                                         unsigned long long *__restrict__ counters, int64_t t,
                                         double *__restrict__ pout = nullptr, const CellIdx *ci = nullptr)
{
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    const CellIdx c0 = ci ? *ci : cell_of(P, t);
    const int i = c0.i0 + 1, j = c0.j0 + 1, k = c0.k0 + 1;
    double g[3][3], gi[3][3], sqg, alpha, beta[3], q[NEQ], c[NEQ];
    metric_c2p<MK>(P, coord_cc(P, i, j), g, gi, sqg, alpha, beta);
    const int64_t pc = pindex(P, i, j, k);
    double w_start = 0.0;
    if (WARM) {
        double v[3] = {p[2 * P.vstride + pc], p[3 * P.vstride + pc], p[4 * P.vstride + pc]};
        double vsq = 0.0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) vsq = vsq + g[a][b] * v[a] * v[b];
        w_start = (vsq < 1.0) ? 1.0 / sqrtr(1.0 - vsq) : 0.0;
    }
#pragma unroll
    for (int v = 0; v < NEQ; v++) c[v] = u[v * ncell + t];
    int brent = 0;
    bool repaired = false;
    const bool atm = c2p(P.g1, P.rho_floor, P.p_floor, g, gi, sqg, c, q, brent, w_start, &repaired);
    if (atm) atomicAdd(&counters[0], 1ull);
    if (brent) atomicAdd(&counters[1], 1ull);
    double *pw = SPLIT ? pout : p;
    // Mixed path: the store addresses are formed here, after the root solve.
    // Computed early, two of them stayed live through it and spilled at 128
    // VGPRs (4 % of the cell kernel's traffic; the kernel is 112 VGPRs
    // without); the FP64 kernel measured 0.3 % faster without the barrier.
    int64_t pcw = pc, tw = t;
    if constexpr (WARM) asm volatile("" : "+v"(pcw), "+v"(tw));
#pragma unroll
    for (int v = 0; v < NEQ; v++) pw[v * P.vstride + pcw] = q[v];
    // U changes only where con2prim rescales the momentum or resets the cell
    if (repaired) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) u[v * ncell + tw] = c[v];
    }
}


// All face fluxes of a stage in one launch: logical blocks [0, nb[0]) are x
// faces, then nb[1] blocks of y faces and nb[2] of z faces.
struct Flux3 {
    int64_t nb[3];
    int kt[3], gw[3];
};
template <int MK, typename R, int SCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_flux3(
    Params P, const double *__restrict__ p, R *__restrict__ Fx, R *__restrict__ Fy, R *__restrict__ Fz,
    int64_t fsx, int64_t fsy, int64_t fsz, Flux3 f)
{
    extern __shared__ double lds_geo[];
    R *lds = reinterpret_cast<R *>(lds_geo);
    int64_t L = xcd_block(blockIdx.x, gridDim.x);
    if (L < f.nb[0]) {
        flux_body<MK, 1, R, SCH>(P, p, Fx, fsx, L, f.kt[0], f.gw[0], lds);
        return;
    }
    L -= f.nb[0];
    if (L < f.nb[1]) {
        flux_body<MK, 2, R, SCH>(P, p, Fy, fsy, L, f.kt[1], f.gw[1], lds);
        return;
    }
    flux_body<MK, 3, R, SCH>(P, p, Fz, fsz, L - f.nb[1], f.kt[2], f.gw[2], lds);
}

// Cell of thread tid of logical block L of the cell kernels (256 cells per
// block), and its indices.  A cell reads the y faces of its row and of the next row, and the z faces of
// its plane and of the next plane.  Blocks walk tiles of 8 rows x 16 planes
// (planes fastest), so the blocks of a tile -- about what one XCD holds at
// once -- share those neighbouring faces in L2; in plain row order the next
// plane's row comes 128 blocks (12 MB of streamed data on the 3D torus) after
// its neighbour and is read from memory again.  Grids whose rows are not
// whole blocks, or whose y and z extents are not whole tiles, keep row order.
// In the tiled walk a block's row and plane are uniform: the divisions are
// done once per wave in 32 bits, not per lane in 64.
constexpr int kCellTileY = 8, kCellTileZ = 16;
__device__ __forceinline__ int64_t cell_block(const Params &P, unsigned L, unsigned tid, CellIdx &ci)
{
    if (P.nx % 256 == 0 && P.ny % kCellTileY == 0 && P.nz % kCellTileZ == 0) {
        const unsigned spr = (unsigned)P.nx / 256u;  // blocks per row
        const unsigned seg = L % spr, r = L / spr;
        const unsigned per = kCellTileY * kCellTileZ, tile = r / per, w = r % per;
        const unsigned ntj = (unsigned)P.ny / kCellTileY;
        ci.j0 = (int)((tile % ntj) * kCellTileY + w / kCellTileZ);
        ci.k0 = (int)((tile / ntj) * kCellTileZ + w % kCellTileZ);
        ci.i0 = (int)(seg * 256u + tid);
        return ((int64_t)ci.k0 * P.ny + ci.j0) * P.nx + ci.i0;
    }
    const int64_t t = (int64_t)L * 256 + tid;
    ci = cell_of(P, t);
    return t;
}

// Cell update and con2prim of the same cell in one thread (stages 1..3).
// Held to 128 VGPRs (four waves per SIMD): the kernel streams about 440 B per
// cell, and the extra wave hides more of that latency than the few spilled
// values cost (FP32 faces: 144 VGPRs and three waves otherwise).  Launched
// with 256-thread blocks (cell_block).
template <int MK, typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_cell(Params P, double *__restrict__ p, const R *__restrict__ Fx,
                                              const R *__restrict__ Fy, const R *__restrict__ Fz,
                                              int64_t fsx, int64_t fsy, int64_t fsz,
                                              const double *__restrict__ u0, double *__restrict__ u,
                                              int stage, double dt, unsigned long long *__restrict__ counters)
{
    CellIdx ci;
    const int64_t t = cell_block(P, xcd_block(blockIdx.x, gridDim.x), threadIdx.x, ci);
    if (t >= (int64_t)P.nx * P.ny * P.nz) return;
    update_body<MK, R>(P, p, Fx, Fy, Fz, fsx, fsy, fsz, u0, u, stage, dt, t, nullptr, nullptr, nullptr, nullptr,
                       &ci);
    c2p_body<MK, sizeof(R) == sizeof(float)>(P, p, u, counters, t, nullptr, &ci);
}

// z-march form of the stage (3D grids): each lane owns one (x, y) column and
// walks a segment of P.zseg cells along z.  It computes the z faces in order
// (the face geometry from the stage's table, as k_flux_tab reads it), keeps
// the last one in registers and differences it with the next, so the z fluxes
// never go through HBM; the x and y fluxes come from their face arrays.  The
// update and con2prim of each cell follow in the same thread, as in k_cell.
// A segment computes its first face, which the previous segment also
// computes (zseg + 1 faces for zseg cells).  The primitives are read from
// pin and written to pout: a later face of the column, or a neighbouring
// segment, still reads the cells this one has updated.  The arithmetic is
// that of k_flux_tab<3> and k_cell, so the results are theirs bit for bit.
template <int MK, typename R, int SCH>
__device__ __forceinline__ void z_face(const Params &P, const double *__restrict__ p, const FaceGeo<R> &fg,
                                       bool zero, int i, int j, int c, R out[NEQ])
{
    if (zero) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) out[v] = 0;
    } else {
        flux_face<MK, 3, R, SCH>(P, p, fg, i, j, c, c, P.nz, out);
    }
}

template <int MK, typename R, int SCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_zcell(
    Params P, const double *__restrict__ pin, double *__restrict__ pout, const R *__restrict__ Fx,
    const R *__restrict__ Fy, int64_t fsx, int64_t fsy, const R *__restrict__ tab, const double *__restrict__ u0,
    double *__restrict__ u, int stage, double dt, unsigned long long *__restrict__ counters)
{
    const int64_t nab = (int64_t)P.nx * P.ny;
    const int64_t nchunk = (nab + 63) / 64;
    // consecutive waves take neighbouring column chunks of one segment, which
    // read the same rows of y faces
    const int64_t w = (int64_t)xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t ab = (w % nchunk) * 64 + (threadIdx.x & 63);
    const int kb = (int)(w / nchunk) * P.zseg + 1;
    if (ab >= nab || kb > P.nz) return;
    const int ke = min(kb + P.zseg, P.nz + 1);
    const int i = (int)(ab % P.nx) + 1, j = (int)(ab / P.nx) + 1;
    const bool zero = (int)tab[FG_FLAGS * nab + ab] & FG_ZERO;
    double *pr = const_cast<double *>(pin);  // read only: c2p_body writes pout
    // the previous face of the column waits in LDS (not in registers, which
    // the face and cell code need in full); each lane touches only its own slots
    __shared__ R zprev[4][NEQ][64];
    R(*fl)[64] = zprev[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    // face c lies between cells c and c + 1; one copy of the face code serves
    // the segment's first face and the others
    for (int c = kb - 1; c < ke; c++) {
        // The face geometry and the centre metric depend on (x, y) only; the
        // loop re-reads them every cell instead of holding them in registers
        // over the loop (the compiler would hoist them, and spill)
        const R *tp = tab + ab;
        const double *cp = P.coord;
        asm volatile("" : "+v"(tp), "+s"(cp));
        Params Pc = P;
        Pc.coord = cp;
        const FaceGeo<R> fg{tp, nab};
        R fr[NEQ];
        z_face<MK, R, SCH>(Pc, pin, fg, zero, i, j, c, fr);
        if (c >= kb) {
            R zl[NEQ];
#pragma unroll
            for (int v = 0; v < NEQ; v++) zl[v] = fl[v][lane];
            const int64_t t = ((int64_t)(c - 1) * P.ny + (j - 1)) * P.nx + (i - 1);
            update_body<MK, R, true>(Pc, pin, Fx, Fy, nullptr, fsx, fsy, 0, u0, u, stage, dt, t, zl, fr);
            c2p_body<MK, sizeof(R) == sizeof(float), true>(Pc, pr, u, counters, t, pout);
        }
#pragma unroll
        for (int v = 0; v < NEQ; v++) fl[v][lane] = fr[v];
    }
}

// x-fused form of the stage (3D grids): the x faces are computed by the cell
// kernel itself.  A block of 256 threads takes 256 consecutive x faces of one
// z plane (flattened over (x, y) as in k_flux_tab), keeps their fluxes in LDS
// and, after a barrier, 255 threads update the cells between consecutive
// faces -- the flux divergence, sources, RK combination and con2prim of
// k_cell -- with the y and z fluxes from their face arrays.  Block b starts
// at face 255 b, so every face pair of the plane (every cell) falls in one
// block and one face in 255 is computed twice; a pair that spans two rows
// (the last face of one row, the first of the next) has no cell.  The x face
// array never goes through HBM, and the cell work (HBM-bound) shares the CUs
// with the face work (VALU-bound) of other blocks.  The primitives are read
// from pin and written to pout: other blocks' stencils still read the cells
// this one updates.  DEFER (scheme 1 with deferred-face lists): the x faces
// through flux_face_hot, and a cell next to a face whose HLLC solve falls
// back to HLLE is appended to defer_list[0] (cell indices; each cell has one
// owner) for k_xcell_fix, which keeps the fallback's code out of this
// kernel.  The arithmetic is that of k_flux_tab<1> and k_cell, so the results
// are theirs bit for bit.
template <int MK, typename R, int SCH, bool DEFER = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_xcell(
    Params P, const double *__restrict__ pin, double *__restrict__ pout, const R *__restrict__ Fy,
    const R *__restrict__ Fz, int64_t fsy, int64_t fsz, const R *__restrict__ tab, const double *__restrict__ u0,
    double *__restrict__ u, int stage, double dt, unsigned long long *__restrict__ counters)
{
    static_assert(!DEFER || SCH == 1, "deferred faces: scheme 1 only");
    __shared__ R fl[NEQ][256];
    __shared__ int okf[256];
    const int nfx = P.nx + 1;
    const int64_t nab = (int64_t)nfx * P.ny;
    const int64_t nchunk = (nab + 253) / 255;
    const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int k0 = (int)(blk / nchunk);
    const int64_t ab = (blk % nchunk) * 255 + threadIdx.x;
    R out[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) out[v] = 0;
    bool ok = true;
    if (ab < nab) {
        const int a = (int)(ab % nfx), b = (int)(ab / nfx);
        const FaceGeo<R> fg{tab + ab, nab};
        if (!(fg.flags() & FG_ZERO)) {
            if constexpr (DEFER) ok = flux_face_hot<MK, 1>(P, pin, fg, a, b + 1, k0 + 1, a, P.nx, out);
            else flux_face<MK, 1, R, SCH>(P, pin, fg, a, b + 1, k0 + 1, a, P.nx, out);
        }
    }
#pragma unroll
    for (int v = 0; v < NEQ; v++) fl[v][threadIdx.x] = out[v];
    if constexpr (DEFER) okf[threadIdx.x] = ok ? 1 : 0;
    __syncthreads();
    if (threadIdx.x == 255 || ab + 1 >= nab) return;
    const int a = (int)(ab % nfx);
    if (a >= P.nx) return;
    const int b = (int)(ab / nfx);
    const int64_t t = ((int64_t)k0 * P.ny + b) * P.nx + a;
    if constexpr (DEFER) {
        const bool fix = !(okf[threadIdx.x] && okf[threadIdx.x + 1]);
        defer_push(P.defer_ctr, P.defer_list[0], (unsigned)((int64_t)nfx * P.ny * P.nz), (unsigned)t, fix);
        if (fix) return;
    }
    R xl[NEQ], xr[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) {
        xl[v] = fl[v][threadIdx.x];
        xr[v] = fl[v][threadIdx.x + 1];
    }
    update_body<MK, R, false, true>(P, pin, nullptr, Fy, Fz, 0, fsy, fsz, u0, u, stage, dt, t, nullptr, nullptr,
                                    xl, xr);
    c2p_body<MK, sizeof(R) == sizeof(float), true>(P, const_cast<double *>(pin), u, counters, t, pout);
}

// The cells k_xcell deferred: each thread takes list entries q = thread,
// thread + grid, ... below the count, computes the cell's two x faces with
// flux_face (HLLC with its HLLE fallback) and runs the cell's update and
// con2prim as k_xcell does.  The working block that finishes last clears the
// count and the block counter (as in k_flux_defer).
template <int MK, typename R, int SCH>
__global__ __launch_bounds__(256) void k_xcell_fix(Params P, const double *__restrict__ pin, double *__restrict__ pout,
                                                   const R *__restrict__ Fy, const R *__restrict__ Fz, int64_t fsy,
                                                   int64_t fsz, const R *__restrict__ tab,
                                                   const double *__restrict__ u0, double *__restrict__ u, int stage,
                                                   double dt, unsigned long long *__restrict__ counters)
{
    unsigned *ctr = P.defer_ctr;
    const unsigned *list = P.defer_list[0];
    const int64_t nab = (int64_t)(P.nx + 1) * P.ny;
    const unsigned n = min(__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), (unsigned)(nab * P.nz));
    const unsigned nbk = min(gridDim.x, (n + blockDim.x - 1) / blockDim.x);  // working blocks, as in k_flux_defer
    if (blockIdx.x >= nbk) return;
    for (unsigned q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += nbk * blockDim.x) {
        const int64_t t = list[q];
        const int a = (int)(t % P.nx);
        const int64_t rem = t / P.nx;
        const int b = (int)(rem % P.ny), k0 = (int)(rem / P.ny);
        R xl[NEQ], xr[NEQ];
        for (int side = 0; side < 2; side++) {
            const int64_t ab = (int64_t)b * (P.nx + 1) + a + side;
            const FaceGeo<R> fg{tab + ab, nab};
            R *o = side ? xr : xl;
#pragma unroll
            for (int v = 0; v < NEQ; v++) o[v] = 0;
            if (!(fg.flags() & FG_ZERO)) flux_face<MK, 1, R, SCH>(P, pin, fg, a + side, b + 1, k0 + 1, a + side, P.nx, o);
        }
        update_body<MK, R, false, true>(P, pin, nullptr, Fy, Fz, 0, fsy, fsz, u0, u, stage, dt, t, nullptr, nullptr,
                                        xl, xr);
        c2p_body<MK, sizeof(R) == sizeof(float), true>(P, const_cast<double *>(pin), u, counters, t, pout);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned fin = __hip_atomic_fetch_add(ctr + 1, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (fin == nbk - 1) {
            __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(ctr + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Blocks of k_xcell.
inline int64_t xcell_blocks(const Params &P)
{
    const int64_t nab = (int64_t)(P.nx + 1) * P.ny;
    return ((nab + 253) / 255) * P.nz;
}

// Stage 1..3 in the x-fused form: y and z faces (as in the split form, into
// their arrays), the x-face geometry into geo_x until it is filled, then
// k_xcell (and k_xcell_fix for its deferred cells).
template <int MK, typename R>
void launch_stage_x_mk(const Params &P, hipStream_t s, const double *pin, double *pout, void *const F[3],
                       const int64_t fstr[3], void *geo_x, const double *u0, double *u, int stage, double dt,
                       unsigned long long *counters)
{
    R *Fy = (R *)F[1], *Fz = (R *)F[2], *tab = (R *)geo_x;
    const int64_t nab = (int64_t)(P.nx + 1) * P.ny;
    const unsigned nb = (unsigned)xcell_blocks(P);
#define GRHD_XSTAGE(SCHV)                                                                                        \
    do {                                                                                                         \
        launch_flux<MK, 2, R, SCHV>(P, s, pin, Fy, fstr[1]);                                                     \
        launch_flux<MK, 3, R, SCHV>(P, s, pin, Fz, fstr[2]);                                                     \
        if (!P.geo_ready) k_geo<MK, 1, R><<<blocks_for(nab, 256), 256, 0, s>>>(P, tab);                          \
        k_xcell<MK, R, SCHV><<<nb, 256, 0, s>>>(P, pin, pout, Fy, Fz, fstr[1], fstr[2], tab, u0, u, stage, dt,   \
                                                counters);                                                       \
    } while (0)
    if constexpr (MK == MK_KS || MK == MK_KS_LOG) {
        const int sch = face_scheme<MK>(P);
        if (sch == 1 && P.defer_list[0]) {
            launch_flux<MK, 2, R, 1>(P, s, pin, Fy, fstr[1]);
            launch_flux<MK, 3, R, 1>(P, s, pin, Fz, fstr[2]);
            if (!P.geo_ready) k_geo<MK, 1, R><<<blocks_for(nab, 256), 256, 0, s>>>(P, tab);
            k_xcell<MK, R, 1, true><<<nb, 256, 0, s>>>(P, pin, pout, Fy, Fz, fstr[1], fstr[2], tab, u0, u, stage, dt,
                                                       counters);
            k_xcell_fix<MK, R, 1><<<kDeferBlocks, 256, 0, s>>>(P, pin, pout, Fy, Fz, fstr[1], fstr[2], tab, u0, u,
                                                               stage, dt, counters);
            return;
        }
        if (sch == 1) { GRHD_XSTAGE(1); return; }
        if (sch == 2) { GRHD_XSTAGE(2); return; }
    }
    GRHD_XSTAGE(0);
#undef GRHD_XSTAGE
}

// Stage 1..3 in the z-march form: x and y faces (as in the split form, into
// their arrays), the z-face geometry into geo_z, then k_zcell.
template <int MK, typename R>
void launch_stage_z_mk(const Params &P, hipStream_t s, const double *pin, double *pout, void *const F[3],
                       const int64_t fstr[3], void *geo_z, const double *u0, double *u, int stage, double dt,
                       unsigned long long *counters)
{
    if (P.zseg == kStageXFused) {
        launch_stage_x_mk<MK, R>(P, s, pin, pout, F, fstr, geo_z, u0, u, stage, dt, counters);
        return;
    }
    R *Fx = (R *)F[0], *Fy = (R *)F[1], *tab = (R *)geo_z;
    const int64_t nab = (int64_t)P.nx * P.ny;
    const int64_t waves = ((nab + 63) / 64) * ((P.nz + P.zseg - 1) / P.zseg);
    const unsigned nb = (unsigned)((waves + 3) / 4);
#define GRHD_ZSTAGE(SCHV)                                                                                        \
    do {                                                                                                         \
        launch_flux<MK, 1, R, SCHV>(P, s, pin, Fx, fstr[0]);                                                     \
        launch_flux<MK, 2, R, SCHV>(P, s, pin, Fy, fstr[1]);                                                     \
        if (!P.geo_ready) k_geo<MK, 3, R><<<blocks_for(nab, 256), 256, 0, s>>>(P, tab);                          \
        k_zcell<MK, R, SCHV><<<nb, 256, 0, s>>>(P, pin, pout, Fx, Fy, fstr[0], fstr[1], tab, u0, u, stage, dt,  \
                                                counters);                                                       \
    } while (0)
    if constexpr (MK == MK_KS || MK == MK_KS_LOG) {
        const int sch = face_scheme<MK>(P);
        if (sch == 1) { GRHD_ZSTAGE(1); return; }
        if (sch == 2) { GRHD_ZSTAGE(2); return; }
    }
    GRHD_ZSTAGE(0);
#undef GRHD_ZSTAGE
}

// fused: one face launch for every direction (for small grids, where launch
// latency dominates, and for grids with an inactive axis, which have two face
// code paths); otherwise one face launch per direction -- measured faster on
// large 3D grids, where the three face code paths share a CU's instruction
// cache.  Either way the flux divergence, RK combination and
// con2prim of a cell run in one thread of one cell kernel.
template <int MK, typename R>
void launch_stage_mk(const Params &P, hipStream_t s, double *p, void *const F[3], const int64_t fstr[3],
                     const double *u0, double *u, int stage, double dt, unsigned long long *counters,
                     bool fused)
{
    const int bs = 256;
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    R *Fx = (R *)F[0], *Fy = (R *)F[1], *Fz = (R *)F[2];
    Flux3 f;
    int groups = 1;
    for (int d = 0; d < 3; d++) {
        const bool on = (d == 0) || (d == 1 && P.ny > 1) || (d == 2 && P.nz > 1);
        // the fused kernel's tiling: 256-thread blocks, LDS geometry
        const int gw = (P.face_gw[d] >= 1 && P.face_gw[d] <= 4) ? P.face_gw[d] : 4;
        const FluxTiling t = flux_tiling(P.nz + (d == 2), gw);
        f.kt[d] = t.kt;
        f.gw[d] = t.gw;
        f.nb[d] = !on ? 0 : (d == 0) ? flux_blocks<1>(P, t) : (d == 1) ? flux_blocks<2>(P, t) : flux_blocks<3>(P, t);
        if (on) groups = std::max(groups, 4 / t.gw);
    }
    const int sch = face_scheme<MK>(P);
    if (fused) {
        const size_t lds = (size_t)groups * NFG * 64 * sizeof(R);
        const unsigned nb = (unsigned)(f.nb[0] + f.nb[1] + f.nb[2]);
        // compiled schemes for FP32 faces only: with FP64 faces the fused
        // kernel spills far more with them (260 B per lane for WENO5 + HLLE
        // against 56 for the run-time scheme) and measured 4 % slower
        if constexpr ((MK == MK_KS || MK == MK_KS_LOG) && sizeof(R) == sizeof(float)) {
            if (sch == 1) k_flux3<MK, R, 1><<<nb, bs, lds, s>>>(P, p, Fx, Fy, Fz, fstr[0], fstr[1], fstr[2], f);
            else if (sch == 2) k_flux3<MK, R, 2><<<nb, bs, lds, s>>>(P, p, Fx, Fy, Fz, fstr[0], fstr[1], fstr[2], f);
            else k_flux3<MK, R, 0><<<nb, bs, lds, s>>>(P, p, Fx, Fy, Fz, fstr[0], fstr[1], fstr[2], f);
        } else {
            k_flux3<MK, R, 0><<<nb, bs, lds, s>>>(P, p, Fx, Fy, Fz, fstr[0], fstr[1], fstr[2], f);
        }
    } else {
        if constexpr (MK == MK_KS || MK == MK_KS_LOG) {
            if (sch == 1) launch_faces<MK, R, 1>(P, s, p, Fx, Fy, Fz, fstr);
            else if (sch == 2) launch_faces<MK, R, 2>(P, s, p, Fx, Fy, Fz, fstr);
            else launch_faces<MK, R, 0>(P, s, p, Fx, Fy, Fz, fstr);
        } else {
            launch_faces<MK, R, 0>(P, s, p, Fx, Fy, Fz, fstr);
        }
    }
    // Update and con2prim in one kernel at every size: u does not go through
    // HBM between them (80 B per cell of traffic saved).
    if (stage > 0) {
        k_cell<MK, R><<<blocks_for(ncell, bs), bs, 0, s>>>(P, p, Fx, Fy, Fz, fstr[0], fstr[1], fstr[2], u0, u,
                                                           stage, dt, counters);
        return;
    }
    // stage 0: the rhs alone (parity entry point)
    k_update<MK, R><<<blocks_for(ncell, bs), bs, 0, s>>>(P, p, Fx, Fy, Fz, fstr[0], fstr[1], fstr[2], u0, u,
                                                     stage, dt);
}

// The stage and face launchers of one metric kind and face arithmetic,
// instantiated by the translation unit of that kind (grhd_stage_*.hip), so
// that the kinds compile in parallel; grhd_stage_dispatch.hip selects them.
#define GRHD_STAGE_UNIT(MKV, R)                                                                          \
    template void launch_stage_mk<MKV, R>(const Params &, hipStream_t, double *, void *const[3],        \
                                          const int64_t[3], const double *, double *, int, double,      \
                                          unsigned long long *, bool);                                  \
    template void launch_face_dir_mk<MKV, R>(const Params &, hipStream_t, const double *, void *const[3], \
                                             const int64_t[3], int);                                    \
    template void launch_stage_z_mk<MKV, R>(const Params &, hipStream_t, const double *, double *,       \
                                            void *const[3], const int64_t[3], void *, const double *,   \
                                            double *, int, double, unsigned long long *);

}  // namespace grhd
