// grhd_ctx.hpp -- the state of one problem on the device (opaque grhd_ctx of
// the C ABI), shared by the single-problem entry points and the batch layer.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/grhd_c.h"
#include "grhd_physics.hpp"

struct grhd_ctx {
    grhd_config cfg;
    grhd::Params P;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool mixed = false;
    bool faces_tuned = false;  // face-kernel tiling chosen (tune_faces)
    int forced_zseg = -1;      // grhd_set_stage_form: -1 timed choice, else P.zseg (kStageXFused: x-fused)
    int forced_face = -1;      // grhd_set_face_form: -1 timed choice, else P.face_gw of every direction
    double *p = nullptr;      // ghosted primitives, SoA [5][NZs][NYs][NXs]
    // second primitive array of the ping-pong stage forms (z-march, P.zseg > 0;
    // x-fused, kStageXFused), which read one and write the other; the two are swapped after each stage,
    // so p is always the current one.  Allocated by the stage-form tuner.
    double *p_alt = nullptr;
    double *u[2] = {nullptr, nullptr};  // conserved, SoA [5][nz][ny][nx]
    int ucur = 0;             // u[ucur] holds U^n
    void *F[3] = {nullptr, nullptr, nullptr};
    void *geo[3] = {nullptr, nullptr, nullptr};  // per-stage face-geometry tables (tuner's table form)
    bool geo_filled[3] = {false, false, false};  // geo[d] holds the problem's face geometry
    int64_t fstr[3] = {0, 0, 0};
    // deferred-face lists and their counters (Params::defer_list, defer_faces)
    unsigned *defer_list[3] = {nullptr, nullptr, nullptr};
    unsigned *defer_ctr = nullptr;
    double *injector = nullptr;
    double *partial = nullptr;
    int partial_n = 0;        // CFL partials (4 doubles each)
    int64_t partial_cap = 0;  // doubles in partial (reductions and observers)
    unsigned long long *counters = nullptr;  // 4 counters, then 4 the tuner's timing runs count into
    double *coord = nullptr;  // coordinate table (Params::coord)
    int64_t steps = 0;
    double dt_cfl = 0.0;
    bool dt_ready = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    int64_t last_launches = 0;
    std::string err;
};

namespace grhd {
// Device bytes grhd_create allocates for cfg (admission by HBM footprint).
uint64_t ctx_footprint(const grhd_config &cfg);
}  // namespace grhd
