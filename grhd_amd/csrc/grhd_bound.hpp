// grhd_bound.hpp -- ghost-zone fill of one problem, as a per-ghost-cell
// device function shared by the single-problem and the batched kernels.
#pragma once
#include "grhd_physics.hpp"

namespace grhd {

// ------------------------------------------------------------- boundaries --
// set_boundary_conditions (conditions.f90:773-1066) restricted to the ghost
// cells the update reads: for each active axis, the ghost layers of the rows
// whose transverse indices are interior.  Every rule there is a gather from
// interior cells (or analytic / cached boundary data), so each ghost cell is
// filled independently.  The outermost layer (index -ng) is never written,
// as in the reference (it keeps its initial value).
GRHD_HD int wrap(int m, int n) { return ((m - 1) % n + n) % n + 1; }

// Number of ghost cells one fill writes (one thread each).
GRHD_HD int64_t ghost_cells(const Params &P)
{
    return 2LL * P.ng * P.ny * P.nz + (P.ny > 1 ? 2LL * P.ng * P.nx * P.nz : 0) +
           (P.nz > 1 ? 2LL * P.ng * P.nx * P.ny : 0);
}

// Ghost cell t (0 <= t < ghost_cells(P)).
template <int MK>
__device__ __forceinline__ void bound_body(const Params &P, double *__restrict__ p,
                                           const double *__restrict__ injector, int64_t t)
{
    const int ng = P.ng, nx = P.nx, ny = P.ny, nz = P.nz;
    const int64_t nxs = 2LL * ng * ny * nz;
    const int64_t nys = (ny > 1) ? 2LL * ng * nx * nz : 0;
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
                r_loc = tab_xc(P, nx + g)[1];  // r (exp of x on logarithmic grids)
                jac = (MK == MK_EF_LOG || MK == MK_KS_LOG || MK == MK_MINK_SPH_LOG) ? r_loc : 1.0;
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
            metric_face<MK>(P, coord_cc(P, ic, gj), m);
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
        } else if (prob == 12 && (hi ? P.open_hi : P.open_lo)) {
            // theta boundary away from the pole (no reference counterpart: the
            // reference requires theta = [0, pi] here, parameters.f90:577-583):
            // outflow copy of the edge cell, rho and p floored, v^theta
            // limited to outflow.  beta^theta = 0 for every metric of this
            // geometry, so the shift-aware limit of the radial rule
            // (conditions.f90:1007-1026) is v^theta = 0.
            load(gi, hi ? ny : 1, gk);
            q[EQ_DE] = rmax(q[EQ_DE], P.rho_floor);
            q[EQ_PR] = rmax(q[EQ_PR], P.p_floor);
            q[EQ_VY] = hi ? rmax(0.0, q[EQ_VY]) : rmin(0.0, q[EQ_VY]);
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

}  // namespace grhd
