// grhd_batch.hip -- many independent problems on one device (C ABI
// grhd_batch_*).  The reference runs one problem per process; this layer
// keeps several resident (each in its own grhd_ctx, admitted by HBM
// footprint) and advances all of them together: every RK3 stage of every
// problem of one metric kind and face arithmetic goes into the same launches
// (grhd_batch_kernels.hpp).  Chunks of steps are captured once into a
// hipGraph and replayed; the launch form and face tiling are chosen by timing
// the candidates on the batch itself (all candidates give bit-identical
// results, so the timed steps are ordinary steps).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/grhd_c.h"
#include "grhd_batch.hpp"
#include "grhd_bound.hpp"
#include "grhd_ctx.hpp"
#include "grhd_kernels.hpp"
#include "grhd_stage.hpp"

using namespace grhd;

namespace {

constexpr int kMaxGroups = 16;  // metric kinds x face arithmetics
constexpr int kMaxChunk = 64;   // steps per captured graph (of the most expensive group)
constexpr int kMaxPace = 8;     // steps of a cheaper group per step of the most expensive one
// A group is paced only when its estimated step is at least this many times
// cheaper than the most expensive group's: near the launch-latency floor the
// estimate cannot separate steps that differ by less, and a cheap group given
// too many steps becomes the chunk's critical path.
constexpr int kMinPace = 3;
constexpr int kTuneSteps = 3;   // steps timed per candidate
// Problems of at most this many cells run whole chunks of steps in one block
// each (kb_small); larger ones use the launch form the autotuner picks.  One
// block is latency-bound (one or two waves per SIMD): measured 10 % faster
// per step than the launch forms at 256 cells, 8 % slower at 512.
constexpr int64_t kSmallCells = 320;

// Launch forms: phase kernels (one face launch per direction, update,
// con2prim, ghosts; or fused: faces, cells, ghosts), groups running
// concurrently; or one persistent kernel per stage and group.  gw = waves
// sharing a face-geometry column.
struct Variant {
    bool persistent;
    bool fused;
    int gw;
};
const Variant kVariants[] = {{false, false, 4}, {false, true, 4}, {false, true, 1}, {true, false, 4},
                             {true, false, 1}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
// The autotuner times the phase forms only.  The persistent stage kernel
// holds the union of the face and cell code at 178 VGPRs (two waves per SIMD
// against four for the phase kernels), runs the run-time face scheme and waits
// at two grid barriers per stage: it measured 3.6-4x slower than the best
// phase form on the 2D and 3D tori, so it stays selectable
// (grhd_batch_set_variant) but is never chosen by timing.
constexpr int kNumTuned = 3;

struct BRun {
    grhd_ctx *ctx = nullptr;
    double t = 0.0;
    int64_t steps = 0;
    bool done = false;
    uint64_t bytes = 0;
};

struct BGroup {
    int mk = 0;
    bool mixed = false;
    bool small = false;
    bool large = false;  // with pacing: problems above kFusedStageCells cells
    std::vector<int> slots;
    int ratio = 1;      // steps per chunk step (pacing)
    int64_t total[NPH] = {0, 0, 0, 0, 0};
    int lds_groups = 1;
    unsigned grid = 0;  // persistent grid
};

}  // namespace

struct grhd_batch {
    hipStream_t stream = nullptr;
    int cap = 0;
    uint64_t budget = 0, used = 0;
    std::vector<BRun> runs;
    RunDesc *d_runs = nullptr, *h_runs = nullptr;
    double *d_dts = nullptr, *h_dts = nullptr;
    int *d_slot = nullptr;
    int64_t *d_begin = nullptr;
    unsigned *d_bar = nullptr, *d_err = nullptr;
    std::vector<BGroup> groups;
    bool dirty = true;
    int forced = 0;     // 0: autotune, else variant index + 1
    int pace = 1;       // at most this many steps of a cheap group per chunk step
    std::vector<int> slot_ratio;   // steps per chunk step of each slot's group
    int variant = 0;    // variant in use (index)
    std::map<std::string, int> tuned;
    struct GraphEntry {
        hipGraphExec_t exec;
        int64_t launches;
    };
    std::map<std::pair<int, int>, GraphEntry> graphs;  // (variant, K)
    int cu = 0;
    uint64_t limit_ticks = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // groups are independent: in the phase form each group's chain of
    // launches runs on its own stream inside the captured graph
    hipStream_t side[kMaxGroups] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kMaxGroups] = {};
    double last_ms = 0.0;
    int64_t last_launches = 0;
    std::string err;
};

namespace {

int bfail(grhd_batch *b, int code, const std::string &m)
{
    if (b) b->err = m;
    return code;
}

#define BCHK(b, call)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return bfail(b, e_ == hipErrorOutOfMemory ? GRHD_ERR_OUT_OF_MEMORY : GRHD_ERR_HIP, \
                         std::string(#call) + ": " + hipGetErrorString(e_));                  \
    } while (0)

// GRHD_TRACE_BATCH=1 prints each HIP call of a chunk (name, sizes, the
// graph's key) to stderr, flushed, before it is made: a crash inside the
// runtime then names the call it happened in.
bool trace_batch()
{
    static const bool on = [] {
        const char *e = std::getenv("GRHD_TRACE_BATCH");
        return e && e[0] == '1';
    }();
    return on;
}
#define BTRACE(...)                                        \
    do {                                                   \
        if (trace_batch()) {                               \
            std::fprintf(stderr, "grhd_batch: " __VA_ARGS__); \
            std::fputc('\n', stderr);                     \
            std::fflush(stderr);                           \
        }                                                  \
    } while (0)

void drop_graphs(grhd_batch *b)
{
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second.exec);
    b->graphs.clear();
}

int64_t blocks_of(int64_t n) { return (n + 255) / 256; }

// Logical blocks of every phase of one problem for face tiling gw.
void run_blocks(const Params &P, int gw, int64_t out[NPH], int &lds_groups)
{
    const int64_t ncell = (int64_t)P.nx * P.ny * P.nz;
    const FluxTiling tx = flux_tiling(P.nz, gw), ty = flux_tiling(P.nz, gw), tz = flux_tiling(P.nz + 1, gw);
    out[PH_FX] = flux_blocks<1>(P, tx);
    out[PH_FY] = P.ny > 1 ? flux_blocks<2>(P, ty) : 0;
    out[PH_FZ] = P.nz > 1 ? flux_blocks<3>(P, tz) : 0;
    out[PH_CELL] = blocks_of(ncell);
    out[PH_BOUND] = blocks_of(ghost_cells(P));
    lds_groups = std::max(lds_groups, 4 / tx.gw);
    if (P.ny > 1) lds_groups = std::max(lds_groups, 4 / ty.gw);
    if (P.nz > 1) lds_groups = std::max(lds_groups, 4 / tz.gw);
}

// Pacing: with pace > 1 and the groups forked onto their own streams, a group
// whose step is cheaper than the most expensive group's takes up to pace steps
// per chunk step, so that small problems (a Michel run of 10^5-10^6 steps) do
// not advance at the step rate of the largest ones.  The step cost is
// estimated as a latency floor plus a per-cell term: about nine launches of
// about 6 us and 0.5 ns per cell (the one-block form of the smallest problems
// measured within 10 % of the launch forms per step, so it takes the same
// floor).  The estimate only sets how many steps a group takes, never what a
// step computes.
void pace_groups(grhd_batch *b, int v)
{
    b->slot_ratio.assign(b->cap, 1);
    double most = 0.0;
    std::vector<double> cost(b->groups.size(), 0.0);
    for (size_t g = 0; g < b->groups.size(); g++) {
        const BGroup &G = b->groups[g];
        int64_t cells = 0;
        for (int sl : G.slots) {
            const Params &P = b->runs[sl].ctx->P;
            cells += (int64_t)P.nx * P.ny * P.nz;
        }
        cost[g] = 54.0e-6 + 0.5e-9 * (double)cells;
        most = std::max(most, cost[g]);
    }
    const bool forked = !kVariants[v].persistent && b->groups.size() > 1;
    for (size_t g = 0; g < b->groups.size(); g++) {
        BGroup &G = b->groups[g];
        G.ratio = 1;
        const int r = (int)(most / cost[g]);
        if (forked && b->pace > 1 && r >= kMinPace) G.ratio = std::min(b->pace, r);
        for (int sl : G.slots) b->slot_ratio[sl] = G.ratio;
    }
}

// Groups and block tables of the unfinished problems for variant v.
int rebuild(grhd_batch *b, int v)
{
    BCHK(b, hipStreamSynchronize(b->stream));
    drop_graphs(b);
    // with pacing, large problems form groups of their own, so that the small
    // problems of the same metric kind are not held to their rate (unless that
    // would exceed the group limit)
    for (int split = (b->pace > 1) ? 1 : 0; split >= 0; split--) {
        b->groups.clear();
        for (int s = 0; s < b->cap; s++) {
            const BRun &r = b->runs[s];
            if (!r.ctx || r.done) continue;
            const int mk = r.ctx->P.mk;
            const bool mixed = r.ctx->mixed;
            const int64_t cells = (int64_t)r.ctx->P.nx * r.ctx->P.ny * r.ctx->P.nz;
            const bool small = cells <= kSmallCells;
            const bool large = split && cells > kFusedStageCells;
            auto it = std::find_if(b->groups.begin(), b->groups.end(), [&](const BGroup &g) {
                return g.mk == mk && g.mixed == mixed && g.small == small && g.large == large;
            });
            if (it == b->groups.end()) {
                b->groups.push_back(BGroup());
                it = b->groups.end() - 1;
                it->mk = mk;
                it->mixed = mixed;
                it->small = small;
                it->large = large;
            }
            it->slots.push_back(s);
        }
        if ((int)b->groups.size() <= kMaxGroups) break;
    }
    if ((int)b->groups.size() > kMaxGroups) return bfail(b, GRHD_ERR_UNSUPPORTED, "too many batch groups");
    const int gw = kVariants[v].gw;
    for (size_t g = 0; g < b->groups.size(); g++) {
        BGroup &G = b->groups[g];
        const int n = (int)G.slots.size();
        std::vector<int64_t> begin((size_t)NPH * (n + 1), 0);
        G.lds_groups = 1;
        for (int r = 0; r < n; r++) {
            int64_t nb[NPH];
            run_blocks(b->runs[G.slots[r]].ctx->P, gw, nb, G.lds_groups);
            for (int ph = 0; ph < NPH; ph++) begin[(size_t)ph * (n + 1) + r + 1] = begin[(size_t)ph * (n + 1) + r] + nb[ph];
        }
        for (int ph = 0; ph < NPH; ph++) G.total[ph] = begin[(size_t)ph * (n + 1) + n];
        BCHK(b, hipMemcpyAsync(b->d_slot + g * b->cap, G.slots.data(), n * sizeof(int), hipMemcpyHostToDevice,
                               b->stream));
        BCHK(b, hipMemcpyAsync(b->d_begin + g * (size_t)NPH * (b->cap + 1), begin.data(),
                               begin.size() * sizeof(int64_t), hipMemcpyHostToDevice, b->stream));
        if (kVariants[v].persistent && !G.small) {
            const size_t lds = (size_t)G.lds_groups * NFG * 64 * (G.mixed ? sizeof(float) : sizeof(double));
            const int occ = G.mixed ? batch_occupancy_float(G.mk, lds) : batch_occupancy_double(G.mk, lds);
            if (occ <= 0) return bfail(b, GRHD_ERR_UNSUPPORTED, "persistent stage kernel does not fit a CU");
            int64_t most = 1;
            for (int ph = 0; ph < NPH; ph++) most = std::max(most, G.total[ph]);
            G.grid = (unsigned)std::min<int64_t>((int64_t)occ * b->cu, most);
        }
    }
    pace_groups(b, v);
    BCHK(b, hipStreamSynchronize(b->stream));
    b->dirty = false;
    b->variant = v;
    return GRHD_OK;
}

BatchTables tables_of(const grhd_batch *b, size_t g, int v)
{
    const BGroup &G = b->groups[g];
    BatchTables T;
    std::memset(&T, 0, sizeof(T));
    T.runs = b->d_runs;
    T.slot = b->d_slot + g * b->cap;
    T.begin = b->d_begin + g * (size_t)NPH * (b->cap + 1);
    T.n = (int)G.slots.size();
    T.gw = kVariants[v].gw;
    T.lds_groups = G.lds_groups;
    for (int ph = 0; ph < NPH; ph++) T.total[ph] = G.total[ph];
    return T;
}

// Enqueue K steps (three stages each) of every group: one chain per group,
// forked onto the side streams in the phase form (the persistent form keeps
// the groups in sequence: they share the grid barrier and the whole device).
void enqueue_steps(grhd_batch *b, int v, int K, int64_t &launches)
{
    const bool fork = !kVariants[v].persistent && b->groups.size() > 1;
    if (fork) {
        (void)hipEventRecord(b->fork_ev, b->stream);
        for (size_t g = 0; g < b->groups.size(); g++) (void)hipStreamWaitEvent(b->side[g], b->fork_ev, 0);
    }
    for (size_t g = 0; g < b->groups.size(); g++) {
        hipStream_t sg = fork ? b->side[g] : b->stream;
        if (b->groups[g].small) {
            BatchLaunch L;
            std::memset(&L, 0, sizeof(L));
            L.tables = tables_of(b, g, v);
            L.dts = b->d_dts;
            L.mk = b->groups[g].mk;
            L.small = true;
            L.K = K * b->groups[g].ratio;
            L.cap = b->cap;
            if (b->groups[g].mixed) batch_stage_float(L, sg);
            else batch_stage_double(L, sg);
            launches += 1;
        }
        for (int s = 0; s < K * b->groups[g].ratio && !b->groups[g].small; s++)
            for (int stage = 1; stage <= 3; stage++) {
                BatchLaunch L;
                std::memset(&L, 0, sizeof(L));
                L.tables = tables_of(b, g, v);
                L.dts = b->d_dts + (size_t)s * b->cap;
                L.sstep = s;
                L.stage = stage;
                L.mk = b->groups[g].mk;
                L.persistent = kVariants[v].persistent;
                L.fused = kVariants[v].fused;
                L.grid = b->groups[g].grid;
                L.bar = b->d_bar;
                L.err = b->d_err;
                L.limit_ticks = b->limit_ticks;
                if (b->groups[g].mixed) batch_stage_float(L, sg);
                else batch_stage_double(L, sg);
                const BatchTables &T = L.tables;
                launches += L.persistent ? 1
                            : L.fused    ? 2 + (T.total[PH_BOUND] > 0)
                                         : (T.total[PH_FX] > 0) + (T.total[PH_FY] > 0) + (T.total[PH_FZ] > 0) + 2 +
                                               (T.total[PH_BOUND] > 0);
            }
        if (fork) {
            (void)hipEventRecord(b->join_ev[g], sg);
            (void)hipStreamWaitEvent(b->stream, b->join_ev[g], 0);
        }
    }
}

// The captured graph of K steps with variant v (captured on first use; the
// tables it reads are rebuilt, and the graphs dropped, when the batch changes).
int graph_for(grhd_batch *b, int v, int K, hipGraphExec_t &exec, int64_t &launches)
{
    const auto key = std::make_pair(v, K);
    auto it = b->graphs.find(key);
    if (it != b->graphs.end()) {
        exec = it->second.exec;
        launches = it->second.launches;
        return GRHD_OK;
    }
    launches = 0;
    hipGraph_t g = nullptr;
    BTRACE("hipStreamBeginCapture (variant %d, %d steps)", v, K);
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    enqueue_steps(b, v, K, launches);
    BTRACE("hipStreamEndCapture");
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess) return bfail(b, GRHD_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    BTRACE("hipGraphInstantiate");
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return bfail(b, GRHD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    b->graphs[key] = {exec, launches};
    return GRHD_OK;
}

// Largest steps-per-chunk-step ratio of the current groups.
int chunk_ratio(const grhd_batch *b)
{
    int m = 1;
    for (const BGroup &G : b->groups) m = std::max(m, G.ratio);
    return m;
}

// dt of every slot for the next K steps of its group's pacing (K times its
// ratio; grhd2.f90:179-184, 229-230: dt from the CFL condition, clamped to the
// final time, t advances first); 0 once a problem has reached its final time
// or taken its steps.  Returns the steps each slot takes.
void fill_dts(grhd_batch *b, int K, std::vector<int> &active)
{
    active.assign(b->cap, 0);
    const int rows = K * chunk_ratio(b);
    for (int s = 0; s < b->cap; s++) {
        BRun &r = b->runs[s];
        double t = r.t;
        const double tf = r.ctx ? r.ctx->cfg.final_time : 0.0;
        const int ks = K * (s < (int)b->slot_ratio.size() ? b->slot_ratio[s] : 1);
        for (int k = 0; k < rows; k++) {
            double dt = 0.0;
            if (k < ks && r.ctx && !r.done && t < tf) {
                dt = r.ctx->dt_cfl;
                if (t + dt > tf) dt = tf - t;
                t = t + dt;
                active[s]++;
            }
            b->h_dts[(size_t)k * b->cap + s] = dt;
        }
    }
}

void fill_runs(grhd_batch *b)
{
    for (int s = 0; s < b->cap; s++) {
        RunDesc &d = b->h_runs[s];
        std::memset(&d, 0, sizeof(d));
        const BRun &r = b->runs[s];
        if (!r.ctx) continue;
        grhd_ctx *c = r.ctx;
        d.P = c->P;
        d.p = c->p;
        d.u[0] = c->u[0];
        d.u[1] = c->u[1];
        for (int k = 0; k < 3; k++) {
            d.F[k] = c->F[k];
            d.fstr[k] = c->fstr[k];
        }
        d.injector = c->injector;
        d.counters = c->counters;
        d.ucur = c->ucur;
    }
}

// Run K steps with variant v; when timed, *ms is the device time of the graph.
int run_chunk(grhd_batch *b, int v, int K, double *ms)
{
    std::vector<int> active;
    fill_dts(b, K, active);
    fill_runs(b);
    const size_t runs_bytes = (size_t)b->cap * sizeof(RunDesc);
    const size_t dts_bytes = (size_t)K * chunk_ratio(b) * b->cap * sizeof(double);
    // both copies stay inside their pinned and device buffers (allocated for
    // kMaxChunk x kMaxPace rows of cap slots)
    if (dts_bytes > (size_t)kMaxChunk * kMaxPace * b->cap * sizeof(double))
        return bfail(b, GRHD_ERR_INVALID_ARGUMENT, "chunk larger than the dt table");
    BTRACE("hipMemcpyAsync runs %zu bytes from %p to %p", runs_bytes, (void *)b->h_runs, (void *)b->d_runs);
    BCHK(b, hipMemcpyAsync(b->d_runs, b->h_runs, runs_bytes, hipMemcpyHostToDevice, b->stream));
    BTRACE("hipMemcpyAsync dts %zu bytes from %p to %p", dts_bytes, (void *)b->h_dts, (void *)b->d_dts);
    BCHK(b, hipMemcpyAsync(b->d_dts, b->h_dts, dts_bytes, hipMemcpyHostToDevice, b->stream));
    hipGraphExec_t exec = nullptr;
    int64_t launches = 0;
    BTRACE("graph for variant %d, %d steps", v, K);
    int rc = graph_for(b, v, K, exec, launches);
    if (rc != GRHD_OK) return rc;
    BTRACE("hipEventRecord ev0");
    BCHK(b, hipEventRecord(b->ev0, b->stream));
    BTRACE("hipGraphLaunch %p (%lld launches)", (void *)exec, (long long)launches);
    BCHK(b, hipGraphLaunch(exec, b->stream));
    BTRACE("hipEventRecord ev1");
    BCHK(b, hipEventRecord(b->ev1, b->stream));
    BTRACE("hipStreamSynchronize");
    BCHK(b, hipStreamSynchronize(b->stream));
    float f = 0.0f;
    BCHK(b, hipEventElapsedTime(&f, b->ev0, b->ev1));
    BTRACE("chunk done, %.3f ms", (double)f);
    if (ms) *ms = f;
    b->last_launches += launches;
    if (kVariants[v].persistent) {
        unsigned e = 0;
        BCHK(b, hipMemcpy(&e, b->d_err, sizeof(e), hipMemcpyDeviceToHost));
        if (e) {
            (void)hipMemset(b->d_err, 0, sizeof(unsigned));
            (void)hipMemset(b->d_bar, 0, 2 * sizeof(unsigned));
            return bfail(b, GRHD_ERR_HIP, "persistent stage: grid barrier timed out (grid not co-resident)");
        }
    }
    for (int s = 0; s < b->cap; s++) {
        BRun &r = b->runs[s];
        if (!r.ctx || r.done || active[s] == 0) {
            if (r.ctx && !r.done && r.t >= r.ctx->cfg.final_time) {
                r.done = true;
                b->dirty = true;
            }
            continue;
        }
        for (int k = 0; k < active[s]; k++) r.t = r.t + b->h_dts[(size_t)k * b->cap + s];
        r.steps += active[s];
        r.ctx->steps += active[s];
        r.ctx->ucur = (r.ctx->ucur + active[s]) & 1;
        if (r.t >= r.ctx->cfg.final_time) {
            r.done = true;
            b->dirty = true;
        }
    }
    return GRHD_OK;
}

std::string signature(const grhd_batch *b)
{
    // groups with their cell-count and z-extent classes: problems of similar
    // shape share a tuned choice (small problems do not depend on it)
    std::string s;
    for (const BGroup &G : b->groups) {
        if (G.small) continue;
        int64_t cells = 0;
        int nzmin = 1 << 30;
        for (int sl : G.slots) {
            const Params &P = b->runs[sl].ctx->P;
            cells += (int64_t)P.nx * P.ny * P.nz;
            nzmin = std::min(nzmin, P.nz);
        }
        s += std::to_string(G.mk) + (G.mixed ? "m" : "d") + ":" + std::to_string((int)std::log2((double)cells + 1)) +
             ":" + std::to_string(std::min(nzmin, 4)) + ";";
    }
    return s;
}

}  // namespace

extern "C" {

int grhd_batch_create(int32_t capacity, uint64_t hbm_budget, grhd_batch **out)
{
    if (!out || capacity < 1 || capacity > 4096) return GRHD_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    grhd_batch *b = new grhd_batch();
    b->cap = capacity;
    b->runs.resize(capacity);
    auto bail = [&](hipError_t e, const char *) {
        const int code = (e == hipErrorOutOfMemory) ? GRHD_ERR_OUT_OF_MEMORY : GRHD_ERR_HIP;
        grhd_batch_destroy(b);
        return code;
    };
    hipError_t e;
    int dev = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return bail(e, "hipGetDevice");
    if ((e = hipDeviceGetAttribute(&b->cu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
        return bail(e, "attribute");
    int khz = 0;
    if ((e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev)) != hipSuccess || khz <= 0)
        khz = 100000;  // 100 MHz
    b->limit_ticks = (uint64_t)khz * 1000ull * 5ull;  // 5 s
    if (hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        if ((e = hipMemGetInfo(&fr, &tot)) != hipSuccess) return bail(e, "hipMemGetInfo");
        hbm_budget = (uint64_t)(0.9 * (double)fr);
    }
    b->budget = hbm_budget;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return bail(e, "stream");
    if ((e = hipMalloc(&b->d_runs, capacity * sizeof(RunDesc))) != hipSuccess) return bail(e, "malloc");
    if ((e = hipHostMalloc(&b->h_runs, capacity * sizeof(RunDesc), hipHostMallocDefault)) != hipSuccess)
        return bail(e, "host malloc");
    if ((e = hipMalloc(&b->d_dts, (size_t)kMaxChunk * kMaxPace * capacity * sizeof(double))) != hipSuccess)
        return bail(e, "malloc");
    if ((e = hipHostMalloc(&b->h_dts, (size_t)kMaxChunk * kMaxPace * capacity * sizeof(double), hipHostMallocDefault)) !=
        hipSuccess)
        return bail(e, "host malloc");
    if ((e = hipMalloc(&b->d_slot, (size_t)kMaxGroups * capacity * sizeof(int))) != hipSuccess)
        return bail(e, "malloc");
    if ((e = hipMalloc(&b->d_begin, (size_t)kMaxGroups * NPH * (capacity + 1) * sizeof(int64_t))) != hipSuccess)
        return bail(e, "malloc");
    if ((e = hipMalloc(&b->d_bar, 4 * sizeof(unsigned))) != hipSuccess) return bail(e, "malloc");
    b->d_err = b->d_bar + 2;
    if ((e = hipMemset(b->d_bar, 0, 4 * sizeof(unsigned))) != hipSuccess) return bail(e, "memset");
    if ((e = hipEventCreate(&b->ev0)) != hipSuccess) return bail(e, "event");
    if ((e = hipEventCreate(&b->ev1)) != hipSuccess) return bail(e, "event");
    if ((e = hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming)) != hipSuccess) return bail(e, "event");
    for (int g = 0; g < kMaxGroups; g++) {
        if ((e = hipStreamCreateWithFlags(&b->side[g], hipStreamNonBlocking)) != hipSuccess) return bail(e, "stream");
        if ((e = hipEventCreateWithFlags(&b->join_ev[g], hipEventDisableTiming)) != hipSuccess)
            return bail(e, "event");
    }
    *out = b;
    return GRHD_OK;
}

int grhd_batch_destroy(grhd_batch *b)
{
    if (!b) return GRHD_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    for (BRun &r : b->runs)
        if (r.ctx) grhd_destroy(r.ctx);
    (void)hipFree(b->d_runs);
    (void)hipHostFree(b->h_runs);
    (void)hipFree(b->d_dts);
    (void)hipHostFree(b->h_dts);
    (void)hipFree(b->d_slot);
    (void)hipFree(b->d_begin);
    (void)hipFree(b->d_bar);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
    for (int g = 0; g < kMaxGroups; g++) {
        if (b->join_ev[g]) (void)hipEventDestroy(b->join_ev[g]);
        if (b->side[g]) (void)hipStreamDestroy(b->side[g]);
    }
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return GRHD_OK;
}

const char *grhd_batch_last_error(const grhd_batch *b) { return b ? b->err.c_str() : ""; }

int grhd_batch_add(grhd_batch *b, const grhd_config *cfg, int32_t *slot)
{
    if (!b || !cfg || !slot) return GRHD_ERR_INVALID_ARGUMENT;
    *slot = -1;
    int s = 0;
    while (s < b->cap && b->runs[s].ctx) s++;
    if (s == b->cap) return bfail(b, GRHD_ERR_CAPACITY, "no free slot in the batch");
    uint64_t bytes = 0;
    int rc = grhd_footprint(cfg, &bytes);
    if (rc != GRHD_OK) return bfail(b, rc, grhd_last_error(nullptr));
    if (b->used + bytes > b->budget) return bfail(b, GRHD_ERR_CAPACITY, "HBM budget of the batch exceeded");
    BCHK(b, hipStreamSynchronize(b->stream));
    grhd_ctx *c = nullptr;
    rc = grhd_create(cfg, &c);
    if (rc != GRHD_OK) return bfail(b, rc, grhd_last_error(nullptr));
    rc = grhd_set_stream(c, b->stream);
    if (rc == GRHD_OK) rc = grhd_initial_conditions(c);
    double dt = 0.0;
    if (rc == GRHD_OK) rc = grhd_compute_dt(c, &dt);
    if (rc != GRHD_OK) {
        b->err = c->err;
        grhd_destroy(c);
        return rc;
    }
    BRun &r = b->runs[s];
    r = BRun();
    r.ctx = c;
    r.bytes = bytes;
    r.done = !(0.0 < cfg->final_time);
    b->used += bytes;
    b->dirty = true;
    *slot = s;
    return GRHD_OK;
}

int grhd_batch_problem(grhd_batch *b, int32_t slot, grhd_ctx **ctx)
{
    if (!b || !ctx || slot < 0 || slot >= b->cap || !b->runs[slot].ctx) return GRHD_ERR_INVALID_ARGUMENT;
    *ctx = b->runs[slot].ctx;
    return GRHD_OK;
}

int grhd_batch_remove(grhd_batch *b, int32_t slot)
{
    if (!b || slot < 0 || slot >= b->cap || !b->runs[slot].ctx) return GRHD_ERR_INVALID_ARGUMENT;
    BCHK(b, hipStreamSynchronize(b->stream));
    BRun &r = b->runs[slot];
    grhd_destroy(r.ctx);
    b->used -= r.bytes;
    r = BRun();
    b->dirty = true;
    return GRHD_OK;
}

int grhd_batch_status(grhd_batch *b, int32_t slot, int32_t *done, double *t, int64_t *steps)
{
    if (!b || slot < 0 || slot >= b->cap || !b->runs[slot].ctx) return GRHD_ERR_INVALID_ARGUMENT;
    const BRun &r = b->runs[slot];
    if (done) *done = r.done ? 1 : 0;
    if (t) *t = r.t;
    if (steps) *steps = r.steps;
    return GRHD_OK;
}

int grhd_batch_set_variant(grhd_batch *b, int32_t variant)
{
    if (!b || variant < 0 || variant > kNumVariants) return GRHD_ERR_INVALID_ARGUMENT;
    b->forced = variant;
    b->dirty = true;
    return GRHD_OK;
}

int grhd_batch_set_pacing(grhd_batch *b, int32_t max_ratio)
{
    if (!b || max_ratio < 1 || max_ratio > kMaxPace) return GRHD_ERR_INVALID_ARGUMENT;
    b->pace = max_ratio;
    b->dirty = true;
    return GRHD_OK;
}

int grhd_batch_info(grhd_batch *b, int32_t *variant, int32_t *active, uint64_t *hbm_used, double *last_ms,
                    int64_t *launches)
{
    if (!b) return GRHD_ERR_INVALID_ARGUMENT;
    int n = 0;
    for (const BRun &r : b->runs) n += (r.ctx && !r.done) ? 1 : 0;
    if (variant) *variant = b->variant + 1;
    if (active) *active = n;
    if (hbm_used) *hbm_used = b->used;
    if (last_ms) *last_ms = b->last_ms;
    if (launches) *launches = b->last_launches;
    return GRHD_OK;
}

int grhd_batch_step(grhd_batch *b, int32_t nsteps)
{
    if (!b || nsteps < 0) return GRHD_ERR_INVALID_ARGUMENT;
    b->last_ms = 0.0;
    b->last_launches = 0;
    // ghost fill of every problem first, as each grhd_step / grhd_advance call
    // does (grhd2.f90:233): problems may have been set or perturbed since
    // (the fill is idempotent on a state that ended with one)
    for (BRun &r : b->runs)
        if (r.ctx) launch_boundaries(r.ctx->P, b->stream, r.ctx->p, r.ctx->injector);
    BCHK(b, hipGetLastError());
    int left = nsteps;
    while (left > 0) {
        bool any = false;
        for (const BRun &r : b->runs) any = any || (r.ctx && !r.done);
        if (!any) break;
        int v = b->forced ? b->forced - 1 : b->variant;
        if (b->dirty) {
            int rc = rebuild(b, v);
            if (rc != GRHD_OK) return rc;
            if (!b->forced) {
                // autotune: time every candidate on kTuneSteps real steps,
                // once per batch shape
                const std::string sig = signature(b);
                auto it = b->tuned.find(sig);
                if (sig.empty()) {
                    // small problems only: no launch form to choose
                } else if (it != b->tuned.end()) {
                    if (it->second != b->variant && (rc = rebuild(b, it->second)) != GRHD_OK) return rc;
                } else if (left >= 2 * kTuneSteps * kNumTuned) {
                    double best = 1e300;
                    int bv = 0;
                    for (int c = 0; c < kNumTuned; c++) {
                        if ((rc = rebuild(b, c)) != GRHD_OK) return rc;
                        double ms = 0.0;
                        // first chunk captures and instantiates, second is timed
                        for (int rep = 0; rep < 2; rep++) {
                            if ((rc = run_chunk(b, c, kTuneSteps, &ms)) != GRHD_OK) return rc;
                            left -= kTuneSteps;
                        }
                        b->last_ms += ms;
                        if (ms < best) {
                            best = ms;
                            bv = c;
                        }
                        if (b->dirty) break;  // a problem finished while tuning
                    }
                    if (!b->dirty) b->tuned[sig] = bv;
                    if ((rc = rebuild(b, bv)) != GRHD_OK) return rc;
                    continue;
                }
            }
            v = b->variant;
        }
        const int K = std::min(left, kMaxChunk);
        double ms = 0.0;
        const int rc = run_chunk(b, v, K, &ms);
        if (rc != GRHD_OK) return rc;
        b->last_ms += ms;
        left -= K;
    }
    return GRHD_OK;
}

}  // extern "C"
