// grhd_batch.hpp -- descriptors shared by the batch host layer (grhd_batch.hip)
// and the batched stage kernels (grhd_batch_kernels.hpp).
#pragma once
#include "grhd_physics.hpp"

namespace grhd {

// One problem of a batch, as the batched kernels read it (device copy).
struct RunDesc {
    Params P;
    double *p;                // ghosted primitives
    double *u[2];             // the two conserved buffers
    void *F[3];               // face fluxes (FP64 or FP32)
    int64_t fstr[3];
    const double *injector;
    unsigned long long *counters;
    int ucur;                 // u[ucur] holds U^n at the first step of the launch chunk
    int pad_;
};

// Phases of one RK stage and their logical-block lists.
enum : int { PH_FX = 0, PH_FY = 1, PH_FZ = 2, PH_CELL = 3, PH_BOUND = 4, NPH = 5 };

struct BatchTables {
    const RunDesc *runs;      // indexed by slot
    const int *slot;          // [n] slots of the group's problems
    const int64_t *begin;     // [NPH][n + 1] first logical block of each problem
    int n;
    int gw;                   // waves sharing one face-geometry column (1, 2 or 4)
    int lds_groups;           // face-geometry groups per block the LDS holds
    int pad_;
    int64_t total[NPH];       // logical blocks per phase
};

struct BatchLaunch {
    BatchTables tables;
    const double *dts;        // dt of every slot at this step (0: skip the problem)
    int sstep;                // step index within the launch chunk
    int stage;                // RK stage 1..3
    int mk;
    bool persistent;
    bool fused;               // phase form: 3 launches per stage instead of 6
    bool small;               // one block per problem runs K whole steps
    int K;                    // steps of the launch (small form)
    int cap;                  // row length of the dts table (small form)
    unsigned grid;            // persistent grid (co-resident blocks)
    unsigned *bar;
    unsigned *err;
    uint64_t limit_ticks;
};

// Instantiated in four units (face arithmetic x metric-kind half) so that
// they compile in parallel; the dispatchers are in grhd_batch.hip.
void batch_stage_double_a(const BatchLaunch &B, hipStream_t s);
void batch_stage_double_b(const BatchLaunch &B, hipStream_t s);
void batch_stage_float_a(const BatchLaunch &B, hipStream_t s);
void batch_stage_float_b(const BatchLaunch &B, hipStream_t s);
int batch_occupancy_double_a(int mk, size_t lds);
int batch_occupancy_double_b(int mk, size_t lds);
int batch_occupancy_float_a(int mk, size_t lds);
int batch_occupancy_float_b(int mk, size_t lds);
inline void batch_stage_double(const BatchLaunch &B, hipStream_t s)
{
    if (B.mk <= 3) batch_stage_double_a(B, s);
    else batch_stage_double_b(B, s);
}
inline void batch_stage_float(const BatchLaunch &B, hipStream_t s)
{
    if (B.mk <= 3) batch_stage_float_a(B, s);
    else batch_stage_float_b(B, s);
}
inline int batch_occupancy_double(int mk, size_t lds)
{
    return mk <= 3 ? batch_occupancy_double_a(mk, lds) : batch_occupancy_double_b(mk, lds);
}
inline int batch_occupancy_float(int mk, size_t lds)
{
    return mk <= 3 ? batch_occupancy_float_a(mk, lds) : batch_occupancy_float_b(mk, lds);
}

}  // namespace grhd
