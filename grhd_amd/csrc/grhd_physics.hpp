// grhd_physics.hpp -- point-wise physics of the GRHD2 update for gfx950.
//
// Everything here is a pure function of its arguments (no global state), usable
// on the device and on the host (initial data).  The metric is evaluated
// analytically where it is needed instead of being read from a stored cache:
// the reference precomputes the same functions into ~1.7 KB/cell of cache
// (initialization.f90:265-391); here each value costs a few flops in registers.
//
// Face quantities (reconstruction, Riemann solver) are templated on the
// arithmetic type R (double for the FP64 path, float for the mixed path);
// con2prim, sources and metric determinants are always double.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>

#include <cmath>
#include <stdint.h>

#define GRHD_HD __host__ __device__ __forceinline__

namespace grhd {

enum : int { EQ_DE = 0, EQ_PR = 1, EQ_VX = 2, EQ_VY = 3, EQ_VZ = 4, NEQ = 5 };

// Metric kinds: metric x geometry x radial map (metrics.f90:94-201).
enum MetricKind : int {
    MK_MINK_CART = 0,
    MK_MINK_CYL = 1,
    MK_MINK_SPH = 2,
    MK_MINK_SPH_LOG = 3,
    MK_EF = 4,
    MK_EF_LOG = 5,
    MK_KS = 6,
    MK_KS_LOG = 7,
};

// Params::face_gw of the geometry-table face form with deferred faces, and
// the flags of the table forms' block orders: z first (kFaceOrderZ) or in
// bands of plane tiles, one band per XCD (kFaceOrderBand), added to 0 or
// kFaceTableDefer.
constexpr int kFaceTableDefer = 16;
constexpr int kFaceOrderZ = 32;
constexpr int kFaceOrderBand = 64;
constexpr int kFaceOrders = kFaceOrderZ | kFaceOrderBand;
// The face forms that read the stage's geometry table, and those that defer
// HLLC fallbacks.
inline constexpr bool face_form_table(int f)
{
    return (f & kFaceOrders) != kFaceOrders && ((f & ~kFaceOrders) == 0 || (f & ~kFaceOrders) == kFaceTableDefer);
}
inline constexpr bool face_form_defers(int f) { return face_form_table(f) && (f & ~kFaceOrders) == kFaceTableDefer; }
// Params::zseg of the x-fused stage form.
constexpr int kStageXFused = -2;

// Problem constants broadcast to every kernel (passed by value).
struct Params {
    int problem;
    int mk;
    int nx, ny, nz, ng;
    int rec, rs, lim, sensor;
    int polar;  // spherical / spheroidal geometry
    int cyl;    // cylindrical geometry
    // fishbone_3d with a theta boundary away from the pole (low / high side):
    // outflow ghost fill there instead of the trans-polar one
    int open_lo, open_hi;
    // waves of a face kernel sharing one LDS copy of the face geometry, per
    // direction (1, 2, 4 or 8), or 0: the geometry evaluated once per stage
    // into geo_tab[d] by its own kernel (no barrier in the face kernel), or
    // kFaceTableDefer: the table form with deferred faces (defer_list); either
    // table form with kFaceOrderZ or kFaceOrderBand added walks its blocks z
    // first or in bands of plane tiles.  The
    // single-problem launchers' tiling, chosen by timing the face kernels on
    // the problem (tune_faces).
    int face_gw[3];
    void *geo_tab[3];
    // 1: the face-geometry tables geo_tab hold this problem's geometry (the
    // metric is static: filled once, when the single-problem tuner timed the
    // table form), so the stage skips the kernels that evaluate them
    int geo_ready;
    // Deferred faces of the table-form face kernels (compiled MP5 + sensor +
    // HLLC scheme): a face whose HLLC solve falls back to HLLE is appended to
    // defer_list[d] (face indices, capacity = the direction's face count) and
    // computed by a second kernel over the list.  defer_ctr[2 d] counts the
    // entries and defer_ctr[2 d + 1] the second kernel's finished blocks; its
    // last block clears both.  Null: every face is computed where it lies.
    unsigned *defer_list[3];
    unsigned *defer_ctr;
    // z-march segment of the single-problem stage: each lane walks this many
    // cells of one (x, y) column, computing the z faces in order and
    // differencing them in registers (k_zcell); 0: separate z-face and cell
    // kernels; kStageXFused: the x faces computed by the cell kernel
    // (k_xcell).  Chosen by timing the stage forms on the problem
    // (tune_faces).
    int zseg;
    double gam, g1, mass, spin;
    double rho_floor, p_floor;
    double x_min, dx, y_min, dy, z_min, dz;
    double inv_dx, inv_dy, inv_dz;
    // device storage of the ghosted primitives: storage index of reference
    // index (i, j, k) is ((k + oz) * NYs + (j + oy)) * NXs + (i + ox)
    int ox, oy, oz;
    int NXs, NYs, NZs;
    int64_t vstride;  // NXs * NYs * NZs
    // boundary data
    double jet_ambient_density, jet_ambient_pressure, jet_ambient_velocity;
    double jet_nozzle_radius, jet_density, jet_pressure, jet_velocity;
    double dust_accretion_constant;
    // coordinate table (coord_table_size doubles, device memory): per radial
    // centre i = -ng..nx+ng: x, r, e^{2x}; per radial face i = 0..nx: x, r,
    // e^{2x}; per theta centre j = -ng..ny+ng: sin, cos; per theta face
    // j = 0..ny: sin, cos
    const double *coord;
};

GRHD_HD int64_t pindex(const Params &P, int i, int j, int k)
{
    return ((int64_t)(k + P.oz) * P.NYs + (j + P.oy)) * P.NXs + (i + P.ox);
}

// Coordinates of the cell-centred grid (initialization.f90:110-115, 191-206).
GRHD_HD double xc(const Params &P, int i) { return P.x_min + ((double)i - 0.5) * P.dx; }
GRHD_HD double xf(const Params &P, int i) { return P.x_min + (double)i * P.dx; }
GRHD_HD double yc(const Params &P, int j) { return P.y_min + ((double)j - 0.5) * P.dy; }
GRHD_HD double yf(const Params &P, int j) { return P.y_min + (double)j * P.dy; }
GRHD_HD double zc(const Params &P, int k) { return P.z_min + ((double)k - 0.5) * P.dz; }

// ----------------------------------------------------------- coordinates --
// A metric evaluation point: x (r, or ln r on logarithmic grids), r, e^{2x},
// sin(theta), cos(theta).  The kernels read these from the coordinate table,
// which the host fills once per problem with the same transcendental
// functions the reference calls (initialization.f90:84-263); everything the
// metric computes from them is evaluated in registers.
struct Coord {
    double x, r, e2, s, c;
};
GRHD_HD constexpr bool log_radius_kind(int mk) { return mk == 3 || mk == 5 || mk == 7; }

// Direct evaluation (host set-up code).
inline Coord coord_of(int mk, double x, double y)
{
    return {x, log_radius_kind(mk) ? std::exp(x) : x, std::exp(2.0 * x), std::sin(y), std::cos(y)};
}

GRHD_HD int64_t coord_table_size(int nx, int ny, int ng)
{
    return 3LL * (nx + 2 * ng + 1) + 3LL * (nx + 1) + 2LL * (ny + 2 * ng + 1) + 2LL * (ny + 1);
}
GRHD_HD const double *tab_xc(const Params &P, int i) { return P.coord + 3 * (i + P.ng); }
GRHD_HD const double *tab_xf(const Params &P, int i) { return P.coord + 3 * (P.nx + 2 * P.ng + 1) + 3 * i; }
GRHD_HD const double *tab_yc(const Params &P, int j)
{
    return P.coord + 3 * (P.nx + 2 * P.ng + 1) + 3 * (P.nx + 1) + 2 * (j + P.ng);
}
GRHD_HD const double *tab_yf(const Params &P, int j)
{
    return P.coord + 3 * (P.nx + 2 * P.ng + 1) + 3 * (P.nx + 1) + 2 * (P.ny + 2 * P.ng + 1) + 2 * j;
}
GRHD_HD Coord coord_at(const double *xr, const double *ys) { return {xr[0], xr[1], xr[2], ys[0], ys[1]}; }
// cell centre (i, j); x face i at centre j; centre i at theta face j
GRHD_HD Coord coord_cc(const Params &P, int i, int j) { return coord_at(tab_xc(P, i), tab_yc(P, j)); }
GRHD_HD Coord coord_fc(const Params &P, int i, int j) { return coord_at(tab_xf(P, i), tab_yc(P, j)); }
GRHD_HD Coord coord_cf(const Params &P, int i, int j) { return coord_at(tab_xc(P, i), tab_yf(P, j)); }

template <typename R> struct Lim;
template <> struct Lim<double> {
    static GRHD_HD double tiny() { return DBL_MIN; }
    static GRHD_HD double eps() { return DBL_EPSILON; }
};
template <> struct Lim<float> {
    static GRHD_HD float tiny() { return FLT_MIN; }
    static GRHD_HD float eps() { return FLT_EPSILON; }
};

// MAX/MIN of the reference.  The hardware min/max (one instruction instead of
// a compare and two selects for FP64) returns the same value for every pair of
// non-NaN operands; only the sign of a zero result may differ when the operands
// are +0 and -0, which no later operation of the path can turn into a different
// non-zero value.
template <typename R> GRHD_HD R rmax(R a, R b) { return fmax(a, b); }
template <typename R> GRHD_HD R rmin(R a, R b) { return fmin(a, b); }
template <typename R> GRHD_HD R sgn1(R a) { return copysign((R)1, a); }
GRHD_HD bool finite(double a) { return isfinite(a); }
GRHD_HD bool finite(float a) { return isfinite(a); }

// FP64 reciprocal and quotient on the device in nine instructions: v_rcp_f64,
// two Newton steps, the remainder correction and v_div_fixup -- the
// operations, in order, of the compiler's IEEE sequence (v_div_scale twice,
// v_rcp_f64, four FMAs, a multiply, an FMA, v_div_fmas, v_div_fixup) without
// the range scaling, which leaves the result unchanged unless the quotient's
// exponent is beyond 2^768 or its operands near the ends of the exponent
// range; no operand of the path comes near (lengths, densities, Lorentz
// factors, metric components).  v_div_fixup gives zero, infinite and NaN
// operands the IEEE results.  The result is the IEEE division's bit for bit.
// The host (initial data, set-up) divides with the IEEE operator.  Used in the cell kernel's metric and
// con2prim (its root solve divides four times per iteration): 7 % fewer
// instructions there and +2.6 % on the FP64 stage of the 3D torus, with the
// final state unchanged bit for bit; and, through rcp and rdiv, in the FP64
// face arithmetic: +1.7 % on the FP64 stage once the face kernels no longer
// carry the HLLE fallback (with it inline they spilled and ran 15 % slower).
// The refinement steps of div64 from an estimate y0 of 1 / b (v_rcp_f64 on
// the device; tests/sqrt_div_check.cpp runs them on the host from perturbed
// estimates), for finite operands: the quotient before v_div_fixup.
GRHD_HD double div_refine(double a, double b, double y0)
{
    const double y1 = fma(fma(-b, y0, 1.0), y0, y0);
    const double y2 = fma(fma(-b, y1, 1.0), y1, y1);
    const double q0 = a * y2;
    return fma(fma(-b, q0, a), y2, q0);
}
GRHD_HD double div64(double a, double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_div_fixup(div_refine(a, b, __builtin_amdgcn_rcp(b)), b, a);
#else
    return a / b;
#endif
}
GRHD_HD double rcp64(double b) { return div64(1.0, b); }

// FP64 square root on the device: v_rsq_f64 and the compiler's own Newton
// sequence for sqrt (two multiplies, seven FMAs, then +-0 and +inf returned
// as given), without its range scaling, which leaves the result unchanged for
// arguments at or above 2^-767 (the scaling only acts below); NaN and
// negative arguments give NaN as before.  No argument of the path comes near
// that bound (velocities, sound speeds, metric determinants, discriminants of
// conserved-state magnitudes).  The result is the IEEE square root's bit for
// bit (tests/sqrt_div_check.cpp).  The host uses the C library's.
// sqrt_refine: the steps from an estimate y of 1 / sqrt(x).
GRHD_HD double sqrt_refine(double x, double y)
{
    double g = x * y, h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}
GRHD_HD double sqrt64(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double g = sqrt_refine(x, __builtin_amdgcn_rsq(x));
    return __builtin_amdgcn_class(x, 0x260) ? x : g;  // +-0, +inf
#else
    return sqrt(x);
#endif
}
// Square root in the face or cell arithmetic: sqrt64 in FP64, the hardware
// instruction in FP32 (the mixed path's units take it with
// -fno-hip-fp32-correctly-rounded-divide-sqrt).
GRHD_HD double sqrtr(double x) { return sqrt64(x); }
GRHD_HD float sqrtr(float x) { return sqrtf(x); }

// a / b, given rb = 1 / b, for divisions that share a divisor or divide by a
// constant.  An IEEE FP64 division costs a reciprocal and about ten dependent
// VALU operations; this costs one multiply and two FMAs and returns the same
// correctly rounded quotient: q = RN(a rb) is within one ulp of a / b, the
// remainder a - b q is exact in an FMA, and RN(q + (a - b q) rb) is RN(a / b)
// when rb = RN(1 / b) (P. Markstein, IBM J. Res. Dev. 34 (1990) 111-119).
// Valid for finite non-zero b and a quotient in the normal range; a zero
// quotient comes back as +0 where the division gives -0.  A non-finite
// quotient stays non-finite (inf may come back as NaN), so finite() tests on
// the result keep their outcome.  FP32 (the mixed path): the same two FMAs
// from the refined hardware reciprocal below.
template <typename R> GRHD_HD R quot(R a, R b, R rb)
{
    const R q = a * rb;
    return fma(fma(-q, b, a), rb, q);
}

// True when x is a constant zero once the metric functions are inlined into
// their caller: the structural zeros of a metric kind (beta^theta, gamma_r
// theta, d_phi g, Gamma^0 entries whose every term vanishes ...).  The sums
// over metric components drop the terms such a zero multiplies, which the
// compiler cannot do by itself (x + 0 y is not x when y is infinite or NaN,
// or x is -0).  For finite operands every sum keeps its value; only the sign
// of an exact zero can differ, as with quot above.
#define GRHD_CZERO(x) (__builtin_constant_p(x) && (x) == 0)

// Reciprocal, division and reciprocal square root in the face arithmetic R.
// FP64: the IEEE results the reference computes, through div64 on the
// device.  FP32 (the mixed path,
// held to an error bound against FP64): the hardware v_rcp_f32 (about 1 ulp)
// refined by one Newton step, y + y (1 - b y), then the remainder correction
// of quot: within about half an ulp, in four instructions instead of the eight
// of the range-scaled division.  Every FP32 operand of the path is a normal
// number (the floors 1e-10 and 1e-13 included), where v_rcp_f32 holds its
// accuracy.
template <typename R> GRHD_HD R rcp(R b)
{
    if constexpr (sizeof(R) == 8) {
        return rcp64(b);
    } else {
#if defined(__HIP_DEVICE_COMPILE__)
        const float y = __builtin_amdgcn_rcpf(b);
        return fmaf(fmaf(-b, y, 1.0f), y, y);
#else
        return 1.0f / b;
#endif
    }
}
template <typename R> GRHD_HD R rdiv(R a, R b)
{
    if constexpr (sizeof(R) == 8) {
        return div64(a, b);
    } else {
#if defined(__HIP_DEVICE_COMPILE__)
        return quot(a, b, __builtin_amdgcn_rcpf(b));
#else
        return a / b;
#endif
    }
}
template <typename R> GRHD_HD R rrsqrt(R x) { return rcp(sqrtr(x)); }

// ---------------------------------------------------------------- metric --
// The 3+1 quantities on a face: lapse, shift, spatial metric, sqrt(det).
struct FaceMetric {
    double alpha, beta[3], g[3][3], det;
};

// Centre quantities: everything sources and con2prim read from the reference
// cache (initialization.f90:283-317).  Only the lambda = 0 Christoffels enter
// the sources (equations.f90:543-548), and the metric depends on (r, theta)
// only, so derivatives are kept for directions 1 and 2.
struct CentreMetric {
    double alpha, beta[3], g[3][3], det, sqg;
    double gm[4][4];
    double ginv[3][3];
    double dlna[4];
    double ch0[4][4];    // Gamma^0_{mu nu}
    double dg[4][4][2];  // d_k g_{mu nu}, k = r (0), theta (1)
};

// Basic metric (alpha, beta, gamma, det) of every kind.  Kerr-Schild:
// metrics.f90:753-838 (physical r), 998-1085 (x = ln r); Eddington-Finkelstein:
// 492-551, 612-670; Minkowski: 209-239, 265-294, 338-368, 419-450.
template <int MK>
GRHD_HD void metric_face(const Params &P, const Coord &q, FaceMetric &m)
{
    const double xp = q.x;
    for (int a = 0; a < 3; a++) {
        m.beta[a] = 0.0;
        for (int b = 0; b < 3; b++) m.g[a][b] = 0.0;
    }
    if (MK == MK_MINK_CART) {
        m.alpha = 1.0;
        m.g[0][0] = m.g[1][1] = m.g[2][2] = 1.0;
        m.det = 1.0;
    } else if (MK == MK_MINK_CYL) {
        m.alpha = 1.0;
        m.g[0][0] = 1.0;
        m.g[1][1] = xp * xp;
        m.g[2][2] = 1.0;
        m.det = xp * xp;
    } else if (MK == MK_MINK_SPH) {
        const double r2 = xp * xp, s = q.s, s2 = s * s;
        m.alpha = 1.0;
        m.g[0][0] = 1.0;
        m.g[1][1] = r2;
        m.g[2][2] = r2 * s2;
        m.det = r2 * r2 * s2;
    } else if (MK == MK_MINK_SPH_LOG) {
        const double r = q.r, r2 = r * r, s = q.s, s2 = s * s;
        m.alpha = 1.0;
        m.g[0][0] = r2;
        m.g[1][1] = r2;
        m.g[2][2] = r2 * s2;
        m.det = r2 * r2 * r2 * s2;
    } else if (MK == MK_EF || MK == MK_EF_LOG) {
        const double M = P.mass;
        const double r = q.r;
        const double s = q.s;
        m.alpha = 1.0 / sqrtr(1.0 + 2.0 * M / r);
        const double g11 = (MK == MK_EF_LOG) ? (1.0 + 2.0 * M / r) * (r * r) : 1.0 + 2.0 * M / r;
        m.beta[0] = (MK == MK_EF_LOG) ? (2.0 * M / r / (1.0 + 2.0 * M / r)) / r
                                      : 2.0 * M / r / (1.0 + 2.0 * M / r);
        const double g22 = r * r, g33 = (r * r) * (s * s);
        m.g[0][0] = g11;
        m.g[1][1] = g22;
        m.g[2][2] = g33;
        m.det = g11 * g22 * g33;
    } else {  // Kerr-Schild
        const double M = P.mass, a = P.spin;
        const double r = q.r;
        const double s = q.s, c = q.c, s2 = s * s, c2 = c * c;
        const double Sig = r * r + (a * a) * c2;
        const double H = div64(M * r, Sig);
        m.alpha = rcp64(sqrtr(1.0 + 2.0 * H));
        if (MK == MK_KS_LOG) {
            m.beta[0] = div64(div64(2.0 * H, 1.0 + 2.0 * H), r);
            m.g[0][0] = (1.0 + 2.0 * H) * (r * r);
            m.g[0][2] = -a * (1.0 + 2.0 * H) * s2 * r;
            m.det = (Sig * Sig * s2 * (1.0 + 2.0 * H)) * (r * r);
        } else {
            m.beta[0] = div64(2.0 * H, 1.0 + 2.0 * H);
            m.g[0][0] = 1.0 + 2.0 * H;
            m.g[0][2] = -a * (1.0 + 2.0 * H) * s2;
            m.det = (Sig * Sig) * s2 * (1.0 + 2.0 * H);
        }
        m.g[2][0] = m.g[0][2];
        m.g[1][1] = Sig;
        m.g[2][2] = (r * r + a * a + 2.0 * H * (a * a) * s2) * s2;
    }
}

// Full centre metric: adds g^{mu nu}, d ln(alpha), d g_{mu nu} (r, theta),
// Gamma^0_{mu nu}, gamma^{ij} by the ADM relation and sqrt(det)
// (initialization.f90:283-317).  ny1: the grid has ny == 1, which zeroes the
// theta derivatives (metrics.f90:309, 385, 609, 745, 919, 1178).
template <int MK>
GRHD_HD void metric_centre(const Params &P, const Coord &q, CentreMetric &m)
{
    const double xp = q.x;
    FaceMetric f;
    metric_face<MK>(P, q, f);
    m.alpha = f.alpha;
    for (int a = 0; a < 3; a++) {
        m.beta[a] = f.beta[a];
        for (int b = 0; b < 3; b++) m.g[a][b] = f.g[a][b];
    }
    m.det = f.det;
    for (int a = 0; a < 4; a++) {
        m.dlna[a] = 0.0;
        for (int b = 0; b < 4; b++) {
            m.gm[a][b] = 0.0;
            m.ch0[a][b] = 0.0;
            m.dg[a][b][0] = 0.0;
            m.dg[a][b][1] = 0.0;
        }
    }
    const bool ny1 = (P.ny == 1);
    if (MK == MK_MINK_CART) {
        m.gm[0][0] = -1.0;
        m.gm[1][1] = m.gm[2][2] = m.gm[3][3] = 1.0;
    } else if (MK == MK_MINK_CYL) {  // metrics.f90:443-481
        m.gm[0][0] = -1.0;
        m.gm[1][1] = 1.0;
        m.gm[2][2] = 1.0 / rmax(xp * xp, 1.0e-20);
        m.gm[3][3] = 1.0;
        m.dg[2][2][0] = 2.0 * xp;
    } else if (MK == MK_MINK_SPH) {  // metrics.f90:287-310
        const double r2 = xp * xp, s = q.s, c = q.c, s2 = s * s;
        m.gm[0][0] = -1.0;
        m.gm[1][1] = 1.0;
        m.gm[2][2] = 1.0 / r2;
        m.gm[3][3] = 1.0 / rmax(r2 * s2, 1.0e-30);
        m.dg[2][2][0] = 2.0 * xp;
        m.dg[3][3][0] = 2.0 * xp * (s * s);
        m.dg[3][3][1] = ny1 ? 0.0 : 2.0 * (xp * xp) * s * c;
    } else if (MK == MK_MINK_SPH_LOG) {  // metrics.f90:361-386
        const double r = q.r, r2 = r * r, s = q.s, c = q.c, s2 = s * s;
        const double e2 = q.e2;
        m.gm[0][0] = -1.0;
        m.gm[1][1] = 1.0 / r2;
        m.gm[2][2] = 1.0 / r2;
        m.gm[3][3] = 1.0 / rmax(r2 * s2, 1.0e-30);
        m.dg[1][1][0] = 2.0 * e2;
        m.dg[2][2][0] = 2.0 * e2;
        m.dg[3][3][0] = 2.0 * e2 * (s * s);
        m.dg[3][3][1] = ny1 ? 0.0 : 2.0 * e2 * s * c;
    } else if (MK == MK_EF || MK == MK_EF_LOG) {  // metrics.f90:525-610, 649-746
        const double M = P.mass;
        const double r = q.r;
        const double s = q.s, c = q.c;
        const double a2 = m.alpha * m.alpha;
        m.gm[0][0] = -1.0 / a2;
        for (int i = 0; i < 3; i++) {
            m.gm[0][i + 1] = m.beta[i] / a2;
            m.gm[i + 1][0] = m.gm[0][i + 1];
        }
        m.gm[1][1] = 1.0 / m.g[0][0] - (m.beta[0] * m.beta[0]) / a2;
        m.gm[2][2] = 1.0 / m.g[1][1] - (m.beta[1] * m.beta[1]) / a2;
        m.gm[3][3] = 1.0 / m.g[2][2] - (m.beta[2] * m.beta[2]) / a2;
        if (MK == MK_EF_LOG) {
            m.dlna[1] = M / (r + 2.0 * M);
            m.ch0[0][0] = 2.0 * (M * M) / (r * r * r);
            m.ch0[0][1] = M * (1.0 + 2.0 * M / r) / r;
            m.ch0[1][0] = m.ch0[0][1];
            m.ch0[1][1] = 2.0 * M * (1.0 + M / r);
            m.ch0[2][2] = -2.0 * M;
            m.ch0[3][3] = m.ch0[2][2] * (s * s);
            m.dg[0][0][0] = -2.0 * M / r;
            m.dg[1][1][0] = 2.0 * (r * r) + 2.0 * M * r;
            m.dg[2][2][0] = 2.0 * (r * r);
            m.dg[3][3][0] = 2.0 * (r * r) * (s * s);
            m.dg[3][3][1] = ny1 ? 0.0 : 2.0 * (r * r) * s * c;
        } else {
            m.dlna[1] = M / (r * r + 2.0 * M * r);
            m.ch0[0][0] = 2.0 * (M * M) / (r * r * r);
            m.ch0[0][1] = M * (1.0 + 2.0 * M / r) / (r * r);
            m.ch0[1][0] = m.ch0[0][1];
            m.ch0[1][1] = 2.0 * M * (1.0 + M / r) / (r * r);
            m.ch0[2][2] = -2.0 * M;
            m.ch0[3][3] = m.ch0[2][2] * (s * s);
            m.dg[0][0][0] = -2.0 * M / (r * r);
            m.dg[0][1][0] = -2.0 * M / (r * r);
            m.dg[1][0][0] = m.dg[0][1][0];
            m.dg[1][1][0] = -2.0 * M / (r * r);
            m.dg[2][2][0] = 2.0 * r;
            m.dg[3][3][0] = 2.0 * r * (s * s);
            m.dg[3][3][1] = ny1 ? 0.0 : (r * r) * 2.0 * s * c;
        }
    } else {  // Kerr-Schild: metrics.f90:800-829 / 1045-1078, derivatives 844-921 / 1091-1180
        const double M = P.mass, a = P.spin;
        const bool lg = (MK == MK_KS_LOG);
        const double r = q.r;
        const double s = q.s, c = q.c, s2 = s * s, c2 = c * c;
        const double Sig = r * r + (a * a) * c2;
        const double H = div64(M * r, Sig);
        const double Del = r * r - 2.0 * M * r + a * a;
        const double rSig = rcp64(Sig), rr = rcp64(r);
        m.gm[0][0] = -(1.0 + 2.0 * H);
        if (lg) {
            m.gm[0][1] = quot(2.0 * H, r, rr);
            m.gm[1][1] = div64(quot(Del, Sig, rSig), r * r);
            m.gm[1][3] = quot(quot(a, Sig, rSig), r, rr);
        } else {
            m.gm[0][1] = 2.0 * H;
            m.gm[1][1] = quot(Del, Sig, rSig);
            m.gm[1][3] = quot(a, Sig, rSig);
        }
        m.gm[1][0] = m.gm[0][1];
        m.gm[3][1] = m.gm[1][3];
        m.gm[2][2] = rSig;
        m.gm[3][3] = rcp64((fabs(s) > 1.0e-15) ? Sig * s2 : Sig * 1.0e-30);
        const double S2 = Sig * Sig, rS2 = rcp64(S2), h2 = 1.0 + 2.0 * H, rh2 = rcp64(h2);
        const double dH_dr = quot(M * ((a * a) * c2 - r * r), S2, rS2);
        const double dH_dth = quot(2.0 * M * r * (a * a) * s * c, S2, rS2);
        m.dlna[1] = lg ? quot(-(dH_dr * r), h2, rh2) : quot(-dH_dr, h2, rh2);
        m.dlna[2] = quot(-dH_dth, h2, rh2);
        double (*dg)[4][2] = m.dg;
        const double dS_dth = -2.0 * (a * a) * s * c;
        if (lg) {
            const double gtr = 2.0 * H, grr = 1.0 + 2.0 * H, grp = -a * (1.0 + 2.0 * H) * s2;
            dg[0][0][0] = r * (2.0 * dH_dr);
            dg[0][1][0] = r * gtr + (r * r) * (2.0 * dH_dr);
            dg[0][3][0] = r * (-2.0 * a * s2 * dH_dr);
            dg[1][1][0] = 2.0 * (r * r) * grr + (r * r * r) * (2.0 * dH_dr);
            dg[1][3][0] = r * grp + (r * r) * (-2.0 * a * s2 * dH_dr);
            dg[2][2][0] = 2.0 * (r * r);
            dg[3][3][0] = r * (2.0 * r * s2 + 2.0 * (a * a) * (s2 * s2) * dH_dr);
            if (!ny1) {
                dg[0][0][1] = 2.0 * dH_dth;
                dg[0][1][1] = r * (2.0 * dH_dth);
                dg[0][3][1] = -2.0 * a * (s2 * dH_dth + 2.0 * H * s * c);
                dg[1][1][1] = (r * r) * (2.0 * dH_dth);
                dg[1][3][1] = r * (-a * ((1.0 + 2.0 * H) * 2.0 * s * c + s2 * 2.0 * dH_dth));
                dg[2][2][1] = dS_dth;
                dg[3][3][1] = 2.0 * (r * r + a * a) * s * c + 2.0 * (a * a) * (s2 * s2) * dH_dth +
                              8.0 * (a * a) * H * s * c * s2;
            }
        } else {
            dg[0][0][0] = 2.0 * dH_dr;
            dg[0][1][0] = 2.0 * dH_dr;
            dg[0][3][0] = -2.0 * a * s2 * dH_dr;
            dg[1][1][0] = 2.0 * dH_dr;
            dg[1][3][0] = -a * s2 * (2.0 * dH_dr);
            dg[2][2][0] = 2.0 * r;
            dg[3][3][0] = 2.0 * r * s2 + 2.0 * (a * a) * (s2 * s2) * dH_dr;
            if (!ny1) {
                dg[0][0][1] = 2.0 * dH_dth;
                dg[0][1][1] = 2.0 * dH_dth;
                dg[0][3][1] = -2.0 * a * (s2 * dH_dth + 2.0 * H * s * c);
                dg[1][1][1] = 2.0 * dH_dth;
                dg[1][3][1] = -a * ((1.0 + 2.0 * H) * 2.0 * s * c + s2 * 2.0 * dH_dth);
                dg[2][2][1] = dS_dth;
                dg[3][3][1] = 2.0 * (r * r + a * a) * s * c + 2.0 * (a * a) * (s2 * s2) * dH_dth +
                              8.0 * (a * a) * H * s * c * s2;
            }
        }
        dg[1][0][0] = dg[0][1][0];
        dg[3][0][0] = dg[0][3][0];
        dg[3][1][0] = dg[1][3][0];
        dg[1][0][1] = dg[0][1][1];
        dg[3][0][1] = dg[0][3][1];
        dg[3][1][1] = dg[1][3][1];
        // Gamma^0_{mu nu} = 1/2 g^{0 rho} (d_mu g_{rho nu} + d_nu g_{rho mu} - d_rho g_{mu nu}),
        // accumulated over rho as in metrics.f90:1203-1241 (lambda = 0 row).
        for (int mu = 0; mu < 4; mu++)
            for (int nu = mu; nu < 4; nu++) {
                double acc = 0.0;
                for (int rho = 0; rho < 4; rho++) {
                    const double dmu = (mu == 1 || mu == 2) ? dg[rho][nu][mu - 1] : 0.0;
                    const double dnu = (nu == 1 || nu == 2) ? dg[rho][mu][nu - 1] : 0.0;
                    const double drho = (rho == 1 || rho == 2) ? dg[mu][nu][rho - 1] : 0.0;
                    if (GRHD_CZERO(m.gm[0][rho])) continue;
                    double dsum;
                    if (GRHD_CZERO(dmu)) dsum = dnu;
                    else if (GRHD_CZERO(dnu)) dsum = dmu;
                    else dsum = dmu + dnu;
                    if (!GRHD_CZERO(drho)) dsum = dsum - drho;
                    if (GRHD_CZERO(dsum)) continue;
                    acc = acc + 0.5 * m.gm[0][rho] * dsum;
                }
                m.ch0[mu][nu] = acc;
                m.ch0[nu][mu] = acc;
            }
    }
    const double a2 = m.alpha * m.alpha, ra2 = rcp64(a2);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            m.ginv[a][b] = (GRHD_CZERO(m.beta[a]) || GRHD_CZERO(m.beta[b]))
                               ? m.gm[a + 1][b + 1]
                               : m.gm[a + 1][b + 1] + quot(m.beta[a] * m.beta[b], a2, ra2);
    m.sqg = sqrtr(m.det);
}

// Lighter centre metric for con2prim and prim_to_cons: gamma_ij, gamma^ij, sqrt(det).
template <int MK>
GRHD_HD void metric_c2p(const Params &P, const Coord &q, double g[3][3], double ginv[3][3],
                        double &sqg, double &alpha, double beta[3])
{
    CentreMetric m;
    metric_centre<MK>(P, q, m);
    for (int a = 0; a < 3; a++) {
        beta[a] = m.beta[a];
        for (int b = 0; b < 3; b++) {
            g[a][b] = m.g[a][b];
            ginv[a][b] = m.ginv[a][b];
        }
    }
    sqg = m.sqg;
    alpha = m.alpha;
}

// ------------------------------------------------------- equations (face) --
// Prelude shared by prim_to_cons, calc_fluxes and the wave speeds:
// v^2 = gamma_ij v^i v^j with the causal clamp to v_max (equations.f90:49-63).
// G is any metric accessor with gam(a, b): Mat3 over a 3x3 array, or the
// face geometry FaceGeo.
template <typename R> struct Mat3 {
    const R (*m)[3];
    GRHD_HD R gam(int a, int b) const { return m[a][b]; }
};

template <typename R, typename G>
GRHD_HD R vsq_clamp(const G &g, R v[3])
{
    const R vmax = (R)(1.0 - 1.0e-5);
    R vsq = 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            if (!GRHD_CZERO(g.gam(a, b))) vsq = vsq + g.gam(a, b) * v[a] * v[b];
    if (vsq >= vmax) {
        v[0] = v[0] * sqrtr(rdiv(vmax, vsq));
        v[1] = v[1] * sqrtr(rdiv(vmax, vsq));
        v[2] = v[2] * sqrtr(rdiv(vmax, vsq));
        vsq = vmax;
    }
    return vsq;
}

template <typename R, typename G>
GRHD_HD void lower(const G &g, const R v[3], R vc[3])
{
    for (int a = 0; a < 3; a++) {
        vc[a] = 0;
        for (int b = 0; b < 3; b++)
            if (!GRHD_CZERO(g.gam(a, b))) vc[a] = vc[a] + g.gam(a, b) * v[b];
    }
}

// prim_to_cons (equations.f90:30-88) / prim_to_cons_face (fluxes.f90:76-122).
template <typename R, typename G>
GRHD_HD void p2c(R g1, const R q[NEQ], const G &g, R sqg, R U[NEQ])
{
    R v[3] = {q[EQ_VX], q[EQ_VY], q[EQ_VZ]}, vc[3];
    const R vsq = vsq_clamp<R>(g, v);
    lower<R>(g, v, vc);
    const R h = (R)1 + rdiv(g1 * q[EQ_PR], q[EQ_DE]);
    const R W = rrsqrt((R)1 - vsq);
    U[EQ_DE] = q[EQ_DE] * W;
    U[EQ_PR] = q[EQ_DE] * h * (W * W) - q[EQ_PR] - q[EQ_DE] * W;
    U[EQ_VX] = q[EQ_DE] * h * (W * W) * vc[0];
    U[EQ_VY] = q[EQ_DE] * h * (W * W) * vc[1];
    U[EQ_VZ] = q[EQ_DE] * h * (W * W) * vc[2];
    for (int e = 0; e < NEQ; e++) U[e] = U[e] * sqg;
}

// ------------------------------------------------------- face geometry --
// Everything the Riemann solver reads about a face that depends on its
// position only: the face metric (alpha, beta^i, gamma_ij, sqrt(gamma)) in the
// face arithmetic, the HLLC orthonormal frame (coframe, gamma^dd, the frame's
// validity and the face speed w = beta^d / (alpha sqrt(gamma^dd))), gamma^dd
// by Cramer's rule for the GRHD wave speeds, and whether the face flux is
// identically zero (cylindrical axis, polar faces).  The metric does not depend
// on phi, so one evaluation serves every phi plane; it lives in LDS, one
// column of NFG values per lane.
enum : int {
    FG_ALPHA = 0, FG_BETA = 1, FG_G = 4, FG_SQG = 10, FG_COF = 11, FG_GDD = 20, FG_W = 21,
    FG_GII = 22, FG_FLAGS = 23, FG_FAC = 24, NFG = 25
};
enum : int { FG_FRAME_OK = 1, FG_ZERO = 2 };
// gamma_ij is symmetric: six stored components 00 01 02 11 12 22.
GRHD_HD constexpr int sym_index(int a, int b)
{
    return (a <= b) ? (a == 0 ? b : (a == 1 ? 2 + b : 5)) : (b == 0 ? a : (b == 1 ? 2 + a : 5));
}

// Structural zeros of the face coframe e^(a)_i (face_frame below) along DIR.
// gamma_{r theta} = gamma_{theta phi} = 0 for every metric kind, so the
// Gram-Schmidt frame has one purely-theta vector (the normal for DIR = 2, the
// first tangent for DIR = 1, the second for DIR = 3) and every other frame
// vector has no theta component: the theta covector has only its i = theta
// component, and the others have none there.
template <int DIR> GRHD_HD constexpr bool cof_zero(int a, int i)
{
    return (a == (DIR == 1 ? 1 : (DIR == 2 ? 0 : 2))) != (i == 1);
}

template <typename R> struct FaceGeo {
    const R *col;          // field f of this face column at col[f * stride]:
    int64_t stride = 64;   // an LDS column (64 lanes) or a per-stage table
    GRHD_HD R get(int f) const { return col[f * stride]; }
    GRHD_HD R alpha() const { return get(FG_ALPHA); }
    GRHD_HD R beta(int d) const { return get(FG_BETA + d); }
    // gamma_{r theta} and gamma_{theta phi} vanish for every metric kind
    GRHD_HD R gam(int a, int b) const
    {
        return (a + b == 1 || a + b == 3) ? (R)0 : get(FG_G + sym_index(a, b));
    }
    GRHD_HD R sqg() const { return get(FG_SQG); }
    GRHD_HD R cof(int a, int i) const { return get(FG_COF + 3 * a + i); }
    GRHD_HD R gdd() const { return get(FG_GDD); }
    GRHD_HD R wface() const { return get(FG_W); }
    GRHD_HD R gii() const { return get(FG_GII); }
    // sqrt(gamma) alpha sqrt(gamma^dd), the factor of local_flux_to_coordinate
    GRHD_HD R fac() const { return get(FG_FAC); }
    GRHD_HD int flags() const { return (int)get(FG_FLAGS); }
};

// calc_fluxes (equations.f90:434-502), DIR = 1..3.
template <typename R, int DIR, typename FG>
GRHD_HD void phys_flux(R g1, const R q[NEQ], const FG &f, R F[NEQ])
{
    R v[3] = {q[EQ_VX], q[EQ_VY], q[EQ_VZ]}, vc[3];
    const R vsq = vsq_clamp<R>(f, v);
    lower<R>(f, v, vc);
    const R vn = v[DIR - 1], bn = f.beta(DIR - 1);
    const R h = (R)1 + rdiv(g1 * q[EQ_PR], q[EQ_DE]);
    const R W = rrsqrt((R)1 - vsq);
    const R veff = f.alpha() * vn - bn;
    F[EQ_DE] = q[EQ_DE] * W * veff;
    F[EQ_PR] = q[EQ_DE] * h * (W * W) * veff + q[EQ_PR] * bn - F[EQ_DE];
    F[EQ_VX] = q[EQ_DE] * h * (W * W) * vc[0] * veff + (DIR == 1 ? f.alpha() * q[EQ_PR] : (R)0);
    F[EQ_VY] = q[EQ_DE] * h * (W * W) * vc[1] * veff + (DIR == 2 ? f.alpha() * q[EQ_PR] : (R)0);
    F[EQ_VZ] = q[EQ_DE] * h * (W * W) * vc[2] * veff + (DIR == 3 ? f.alpha() * q[EQ_PR] : (R)0);
    for (int e = 0; e < NEQ; e++) F[e] = F[e] * f.sqg();
}

// Characteristic speeds: SRHD for Minkowski+Cartesian (fluxes.f90:669-720),
// GRHD otherwise (fluxes.f90:724-788; the n_steps==0 diagnostic stop is not
// reproduced).
template <typename R, int DIR, bool SRHD, typename FG>
GRHD_HD void wave_speeds(R g1, R gam, const R q[NEQ], const FG &f, R &wmin, R &wmax)
{
    R v[3] = {q[EQ_VX], q[EQ_VY], q[EQ_VZ]};
    const R vsq = vsq_clamp<R>(f, v);
    const R enth = (R)1 + rdiv(g1 * q[EQ_PR], q[EQ_DE]);
    const R cs2 = rdiv(gam * q[EQ_PR], q[EQ_DE] * enth);
    const R cs = sqrtr(cs2);
    const R vn = v[DIR - 1];
    R e1, e2, e3;
    if (SRHD) {
        const R t1 = rcp((R)1 - vsq * cs2);
        const R t2 = cs * sqrtr(((R)1 - vsq) * (rcp(t1) - (vn * vn) * ((R)1 - cs2)));
        e1 = vn;
        e2 = t1 * (vn * ((R)1 - cs2) + t2);
        e3 = t1 * (vn * ((R)1 - cs2) - t2);
    } else {
        const R gii = f.gii();
        const R bn = f.beta(DIR - 1);
        const R t1 = rdiv(f.alpha(), (R)1 - vsq * cs2);
        const R t2 = cs * sqrtr(rmax((R)0, ((R)1 - vsq) * (gii * ((R)1 - vsq * cs2) -
                                                        (vn * vn) * ((R)1 - cs2))));
        e1 = f.alpha() * vn - bn;
        e2 = t1 * (vn * ((R)1 - cs2) + t2) - bn;
        e3 = t1 * (vn * ((R)1 - cs2) - t2) - bn;
    }
    wmax = rmax(rmax(rmax((R)0, e1), e2), e3);
    wmin = rmin(rmin(rmin((R)0, e1), e2), e3);
}

// calc_hlle_fluxes (fluxes.f90:125-156).
template <typename R, int DIR, bool SRHD, typename FG>
GRHD_HD void hlle(R g1, R gam, const R qL[NEQ], const R qR[NEQ], const FG &f, R F[NEQ])
{
    R fL[NEQ], fR[NEQ], uL[NEQ], uR[NEQ], lLmin, lLmax, lRmin, lRmax;
    phys_flux<R, DIR>(g1, qL, f, fL);
    phys_flux<R, DIR>(g1, qR, f, fR);
    p2c<R>(g1, qL, f, f.sqg(), uL);
    p2c<R>(g1, qR, f, f.sqg(), uR);
    wave_speeds<R, DIR, SRHD>(g1, gam, qL, f, lLmin, lLmax);
    wave_speeds<R, DIR, SRHD>(g1, gam, qR, f, lRmin, lRmax);
    const R ap = rmax(rmax((R)0, lLmax), lRmax);
    const R am = rmin(rmin((R)0, lLmin), lRmin);
    if (fabs(ap - am) < (R)1.0e-14) {
        for (int e = 0; e < NEQ; e++) F[e] = 0;
    } else {
        const R den = ap - am, rden = rcp(den);
        for (int e = 0; e < NEQ; e++)
            F[e] = quot(ap * fL[e] - am * fR[e] + ap * am * (uR[e] - uL[e]), den, rden);
    }
}

// ---------------------------------------------------------------- HLLC --
template <typename R>
GRHD_HD R metric_dot(const R a[3], const R b[3], const R g[3][3])
{
    R s = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) s = s + g[i][j] * a[i] * b[j];
    return s;
}

// build_face_orthonormal_frame (fluxes.f90:402-473): e_(1) normal to the
// face, Gram-Schmidt in the gamma_ij inner product.  Returns the coframe
// cof[a][i] = e^(a)_i and gamma^dd.
template <typename R, int DIR>
GRHD_HD bool face_frame(const R g[3][3], R cof[3][3], R &gdd)
{
    R ginv[3][3], frame[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            cof[i][j] = 0;
            frame[i][j] = 0;
            if (!finite(g[i][j])) return false;
        }
    gdd = 0;
    const R det = g[0][0] * (g[1][1] * g[2][2] - g[1][2] * g[2][1]) -
                  g[0][1] * (g[1][0] * g[2][2] - g[1][2] * g[2][0]) +
                  g[0][2] * (g[1][0] * g[2][1] - g[1][1] * g[2][0]);
    if (!finite(det) || det <= (R)1.0e-30) return false;
    const R rdet = rcp(det);
    ginv[0][0] = quot(g[1][1] * g[2][2] - g[1][2] * g[2][1], det, rdet);
    ginv[0][1] = quot(g[0][2] * g[2][1] - g[0][1] * g[2][2], det, rdet);
    ginv[0][2] = quot(g[0][1] * g[1][2] - g[0][2] * g[1][1], det, rdet);
    ginv[1][0] = quot(g[1][2] * g[2][0] - g[1][0] * g[2][2], det, rdet);
    ginv[1][1] = quot(g[0][0] * g[2][2] - g[0][2] * g[2][0], det, rdet);
    ginv[1][2] = quot(g[0][2] * g[1][0] - g[0][0] * g[1][2], det, rdet);
    ginv[2][0] = quot(g[1][0] * g[2][1] - g[1][1] * g[2][0], det, rdet);
    ginv[2][1] = quot(g[0][1] * g[2][0] - g[0][0] * g[2][1], det, rdet);
    ginv[2][2] = quot(g[0][0] * g[1][1] - g[0][1] * g[1][0], det, rdet);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (!finite(ginv[i][j])) return false;
    const int d = DIR - 1;
    gdd = ginv[d][d];
    if (gdd <= (R)1.0e-30 || !finite(gdd)) return false;
    const R sgdd = sqrtr(gdd), rsgdd = rcp(sgdd);
    for (int i = 0; i < 3; i++) frame[i][0] = quot(ginv[i][d], sgdd, rsgdd);
    const int t0 = (DIR == 1) ? 1 : (DIR == 2) ? 2 : 0;
    const int t1 = (DIR == 1) ? 2 : (DIR == 2) ? 0 : 1;
    for (int b = 1; b < 3; b++) {
        R seed[3] = {0, 0, 0};
        seed[b == 1 ? t0 : t1] = 1;
        for (int a = 0; a < b; a++) {
            const R fa[3] = {frame[0][a], frame[1][a], frame[2][a]};
            const R proj = metric_dot(seed, fa, g);
            for (int i = 0; i < 3; i++) seed[i] = seed[i] - proj * fa[i];
        }
        const R nsq = metric_dot(seed, seed, g);
        if (!finite(nsq) || nsq <= (R)1.0e-28) return false;
        const R snsq = sqrtr(nsq), rsnsq = rcp(snsq);
        for (int i = 0; i < 3; i++) frame[i][b] = quot(seed[i], snsq, rsnsq);
    }
    for (int a = 0; a < 3; a++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) cof[a][i] = cof[a][i] + g[i][j] * frame[j][a];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (!finite(frame[i][j]) || !finite(cof[i][j])) return false;
    return true;
}

// The HLLC helpers below evaluate every quantity of the reference's path
// unconditionally, in the reference's operations and order, and each of its
// fallback tests clears a validity flag instead of returning.  A rejected
// state may then compute values from a degenerate input (infinities, NaN; no
// traps on the device), which the flag discards; an accepted one gets the
// branching form's result bit for bit.  Each helper is one basic block the
// compiler schedules across instead of up to eight short ones: +5 % on the
// FP64 stage of the 3D torus.
//
// hllc_flat, the solver itself in the same form (used for FP32 faces: SALU
// instructions per wave halve and the mixed stage measured 2 % faster).  FP64
// faces keep the solver's own early returns (hllc_try): the flat solver holds
// both sides' star states and more at once, spills 400 B per lane at 128
// VGPRs and halved the FP64 stage rate.

// calc_srhd_local_state (fluxes.f90:509-568): U, F and the speeds of the
// local state; returns the reference's validity.
template <typename R>
GRHD_HD bool srhd_local(R g1, R gam, const R ql[NEQ], R U[NEQ], R F[NEQ], R &wmin, R &wmax)
{
    const R vmax = (R)(1.0 - 1.0e-5);
    const R rho = ql[EQ_DE], pr = ql[EQ_PR];
    const R v0 = ql[EQ_VX], v1 = ql[EQ_VY], v2 = ql[EQ_VZ];
    bool ok = finite(rho) & finite(pr) & finite(v0) & finite(v1) & finite(v2) & !(rho <= (R)0) & !(pr < (R)0);
    const R vsq = v0 * v0 + v1 * v1 + v2 * v2;
    ok = ok & finite(vsq) & !(vsq < (R)0) & !(vsq >= vmax);
    const R h = (R)1 + rdiv(g1 * pr, rho);
    ok = ok & finite(h) & !(h <= (R)0);
    const R cs2 = rdiv(gam * pr, rho * h);
    ok = ok & finite(cs2) & !(cs2 < (R)0) & !(cs2 >= (R)1);
    const R W = rrsqrt((R)1 - vsq);
    const R D = rho * W;
    const R tau = rho * h * (W * W) - pr - D;
    const R m0 = rho * h * (W * W) * v0, m1 = rho * h * (W * W) * v1, m2 = rho * h * (W * W) * v2;
    U[EQ_DE] = D;
    U[EQ_PR] = tau;
    U[EQ_VX] = m0;
    U[EQ_VY] = m1;
    U[EQ_VZ] = m2;
    F[EQ_DE] = D * v0;
    F[EQ_PR] = m0 - D * v0;
    F[EQ_VX] = m0 * v0 + pr;
    F[EQ_VY] = m1 * v0;
    F[EQ_VZ] = m2 * v0;
    const R den = (R)1 - vsq * cs2;
    const R rad = ((R)1 - vsq) * ((R)1 - vsq * cs2 - (v0 * v0) * ((R)1 - cs2));
    ok = ok & !(den <= (R)1.0e-14) & !(rad < (R)-1.0e-12);
    const R ac = sqrtr(cs2) * sqrtr(rmax((R)0, rad));
    const R rden = rcp(den);
    wmin = quot(v0 * ((R)1 - cs2) - ac, den, rden);
    wmax = quot(v0 * ((R)1 - cs2) + ac, den, rden);
#pragma unroll
    for (int e = 0; e < NEQ; e++) ok = ok & finite(U[e]) & finite(F[e]);
    return ok & finite(wmin) & finite(wmax);
}

// conservative_state_is_admissible (fluxes.f90:571-591).
template <typename R>
GRHD_HD bool admissible(const R U[NEQ])
{
    bool ok = true;
#pragma unroll
    for (int e = 0; e < NEQ; e++) ok = ok & finite(U[e]);
    const R E = U[EQ_PR] + U[EQ_DE];
    const R msq = U[EQ_VX] * U[EQ_VX] + U[EQ_VY] * U[EQ_VY] + U[EQ_VZ] * U[EQ_VZ];
    const R scale = rmax(rmax(Lim<R>::tiny(), E * E), U[EQ_DE] * U[EQ_DE] + msq);
    return ok & !(U[EQ_DE] <= (R)0) & !(E <= (R)0) & !(E * E - U[EQ_DE] * U[EQ_DE] - msq <= (R)1.0e-13 * scale);
}

// local_flux_to_coordinate (fluxes.f90:596-624).  The factor
// sqrt(gamma) alpha sqrt(gamma^dd) depends on the face's position only and
// comes with its geometry.
template <int DIR, typename R, typename FG>
GRHD_HD bool local_to_coord(const R Ul[NEQ], const R Fl[NEQ], R w, const FG &fg, R F[NEQ])
{
    R mf[NEQ];
    bool ok = true;
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        mf[e] = Fl[e] - w * Ul[e];
        ok = ok & finite(mf[e]);
    }
    const R fac = fg.fac();  // sqg * alpha * sqrt(gdd), from the geometry
    ok = ok & finite(fac);
    F[EQ_DE] = fac * mf[EQ_DE];
    F[EQ_PR] = fac * mf[EQ_PR];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        F[EQ_VX + i] = 0;
#pragma unroll
        for (int a = 0; a < 3; a++)
            if (!cof_zero<DIR>(a, i)) F[EQ_VX + i] = F[EQ_VX + i] + fac * fg.cof(a, i) * mf[EQ_VX + a];
    }
#pragma unroll
    for (int e = 0; e < NEQ; e++) ok = ok & finite(F[e]);
    return ok;
}

// calc_hllc_fluxes (fluxes.f90:174-397).  Returns false where the reference
// falls back to HLLE; the caller then evaluates HLLE.
template <typename R, int DIR, typename FG>
GRHD_HD bool hllc_flat(R g1, R gam, const R qL[NEQ], const R qR[NEQ], const FG &f, R F[NEQ])
{
    R qLh[NEQ], qRh[NEQ], uL[NEQ], uR[NEQ], fL[NEQ], fR[NEQ];
    R lLmin, lLmax, lRmin, lRmax;
    // lapse, sqrt(gamma) and the face frame were checked with the geometry
    // (fluxes.f90:195-205, 402-473)
    bool ok = (f.flags() & FG_FRAME_OK) != 0;
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        qLh[e] = qL[e];
        qRh[e] = qR[e];
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        qLh[EQ_VX + a] = 0;
        qRh[EQ_VX + a] = 0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (cof_zero<DIR>(a, i)) continue;
            qLh[EQ_VX + a] = qLh[EQ_VX + a] + f.cof(a, i) * qL[EQ_VX + i];
            qRh[EQ_VX + a] = qRh[EQ_VX + a] + f.cof(a, i) * qR[EQ_VX + i];
        }
    }
    ok = ok & srhd_local(g1, gam, qLh, uL, fL, lLmin, lLmax);
    ok = ok & srhd_local(g1, gam, qRh, uR, fR, lRmin, lRmax);
    const R ll = rmin(lLmin, lRmin), lr = rmax(lLmax, lRmax);
    const R w = f.wface();
    ok = ok & finite(w) & !(lr - ll <= (R)1.0e-13);
    // supersonic faces take an outer state; the star region is then not used
    const bool outL = (w <= ll), outR = !outL & (w >= lr);
    R uh[NEQ], fh[NEQ];
    const R dhll = lr - ll;
    const R rhll = rcp(dhll);
    bool okS = true;
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        uh[e] = quot(lr * uR[e] - ll * uL[e] + fL[e] - fR[e], dhll, rhll);
        fh[e] = quot(lr * fL[e] - ll * fR[e] + lr * ll * (uR[e] - uL[e]), dhll, rhll);
        okS = okS & finite(uh[e]) & finite(fh[e]);
    }
    const R Eh = uh[EQ_PR] + uh[EQ_DE], FEh = fh[EQ_PR] + fh[EQ_DE];
    const R A = FEh, B = -(Eh + fh[EQ_VX]), Cq = uh[EQ_VX];
    const R disc0 = B * B - (R)4 * A * Cq;
    const R dscale = rmax(rmax(Lim<R>::tiny(), B * B), fabs((R)4 * A * Cq));
    okS = okS & finite(disc0) & !(disc0 < (R)-1.0e-12 * dscale);
    const R disc = rmax((R)0, disc0);
    const R cscale = rmax(rmax(Lim<R>::tiny(), fabs(B)), fabs(Cq));
    const bool quad = fabs(A) > (R)1.0e-13 * cscale;
    // the root lambda*: linear equation, or the cancellation-free root of the
    // quadratic, or its alternate form; one division of the chosen operands
    const R sd = sqrtr(disc);
    const R dthr = (R)1.0e-14 * rmax(rmax(Lim<R>::tiny(), fabs(B)), sd);
    const R den1 = -B - sd, den2 = -B + sd;
    const bool first = fabs(den1) > dthr;
    okS = okS & (quad ? (first | !(fabs(den2) <= dthr))
                      : !(fabs(B) <= (R)1.0e-14 * rmax(Lim<R>::tiny(), fabs(Cq))));
    const R lnum = !quad ? -Cq : first ? den1 : (R)2 * Cq;
    const R lden = !quad ? B : first ? (R)2 * A : den2;
    R ls = rdiv(lnum, lden);
    const R rtol = (R)1.0e-10 * rmax(rmax((R)1, fabs(ll)), fabs(lr));
    bool valid = finite(ls) & (ls >= ll - rtol) & (ls <= lr + rtol) & (fabs(ls) <= (R)1 + rtol);
    const R ra = rdiv(den2, (R)2 * A);
    const bool alt = !valid & quad & finite(ra) & (ra >= ll - rtol) & (ra <= lr + rtol) & (fabs(ra) <= (R)1 + rtol);
    ls = alt ? ra : ls;
    okS = okS & (valid | alt);
    ls = rmin(lr, rmax(ll, ls));
    const R AL = ll * (uL[EQ_PR] + uL[EQ_DE]) - uL[EQ_VX];
    const R BL = uL[EQ_VX] * (ll - qLh[EQ_VX]) - qLh[EQ_PR];
    const R AR = lr * (uR[EQ_PR] + uR[EQ_DE]) - uR[EQ_VX];
    const R BR = uR[EQ_VX] * (lr - qRh[EQ_VX]) - qRh[EQ_PR];
    const R denL = (R)1 - ll * ls, denR = (R)1 - lr * ls;
    okS = okS & !(fabs(denL) <= (R)1.0e-13) & !(fabs(denR) <= (R)1.0e-13);
    const R psL = rdiv(AL * ls - BL, denL);
    const R psR = rdiv(AR * ls - BR, denR);
    const R sscale = rmax(rmax(rmax(rmax(Lim<R>::tiny(), fabs(uL[EQ_PR] + uL[EQ_DE])),
                                    fabs(uR[EQ_PR] + uR[EQ_DE])),
                               fabs(uL[EQ_VX])),
                          fabs(uR[EQ_VX]));
    const R ptol = (R)1.0e-8 * rmax(rmax(rmax(rmax(Lim<R>::tiny(), fabs(qLh[EQ_PR])),
                                              fabs(qRh[EQ_PR])),
                                         fabs(psL)),
                                    fabs(psR)) +
                   (R)1.0e2 * Lim<R>::eps() * sscale;
    okS = okS & finite(psL) & finite(psR) & !(fabs(psL - psR) > ptol);
    const R ps = (R)0.5 * (psL + psR);
    okS = okS & finite(ps) & !(fabs(ll - ls) <= (R)1.0e-13) & !(fabs(lr - ls) <= (R)1.0e-13);
    // star states (Mignone-Bodo): both must be admissible with finite fluxes
    // (fluxes.f90:351-380); the one on the side of the face speed is used.
    // The other side is checked first and dropped, which keeps one star state
    // live instead of two.
    const bool starL = (w <= ls);
    bool okO = true;
    {
        const R lo = starL ? lr : ll, dso = lo - ls, rdso = rcp(dso);
        R uo[NEQ];
#pragma unroll
        for (int e = 0; e < NEQ; e++) {
            const R uk = starL ? uR[e] : uL[e], fk = starL ? fR[e] : fL[e];
            const R pv = (e == EQ_PR) ? ls : (e == EQ_VX) ? (R)1 : (R)0;
            // pv vanishes for D and the transverse momenta: no pressure term there
            uo[e] = quot((e == EQ_PR || e == EQ_VX) ? lo * uk - fk + ps * pv : lo * uk - fk, dso, rdso);
            okO = okO & finite(fk + lo * (uo[e] - uk));
        }
        okO = okO & admissible(uo);
    }
    const R lk = starL ? ll : lr, dsk = lk - ls, rdsk = rcp(dsk);
    R us[NEQ], fs[NEQ];
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        const R uk = starL ? uL[e] : uR[e], fk = starL ? fL[e] : fR[e];
        const R pv = (e == EQ_PR) ? ls : (e == EQ_VX) ? (R)1 : (R)0;
        us[e] = quot((e == EQ_PR || e == EQ_VX) ? lk * uk - fk + ps * pv : lk * uk - fk, dsk, rdsk);
        fs[e] = fk + lk * (us[e] - uk);
        okS = okS & finite(fs[e]);
    }
    okS = okS & okO & admissible(us);
    // the state the face speed falls in, back to coordinates
    R Us[NEQ], Fs[NEQ];
#pragma unroll
    for (int e = 0; e < NEQ; e++) {
        Us[e] = outL ? uL[e] : outR ? uR[e] : us[e];
        Fs[e] = outL ? fL[e] : outR ? fR[e] : fs[e];
    }
    const bool okT = local_to_coord<DIR>(Us, Fs, w, f, F);
    return ok & (outL | outR | okS) & okT;
}

// calc_hllc_fluxes (fluxes.f90:174-397).  Returns false where the reference
// falls back to HLLE; the caller then evaluates HLLE.
template <typename R, int DIR, typename FG>
GRHD_HD bool hllc_try(R g1, R gam, const R qL[NEQ], const R qR[NEQ], const FG &f, R F[NEQ])
{
    if constexpr (sizeof(R) == sizeof(float)) return hllc_flat<R, DIR>(g1, gam, qL, qR, f, F);
    R qLh[NEQ], qRh[NEQ], uL[NEQ], uR[NEQ], fL[NEQ], fR[NEQ];
    R lLmin, lLmax, lRmin, lRmax;
    // lapse, sqrt(gamma) and the face frame were checked with the geometry
    // (fluxes.f90:195-205, 402-473)
    if (!(f.flags() & FG_FRAME_OK)) return false;
    for (int e = 0; e < NEQ; e++) {
        qLh[e] = qL[e];
        qRh[e] = qR[e];
    }
    for (int a = 0; a < 3; a++) {
        qLh[EQ_VX + a] = 0;
        qRh[EQ_VX + a] = 0;
        for (int i = 0; i < 3; i++) {
            if (cof_zero<DIR>(a, i)) continue;
            qLh[EQ_VX + a] = qLh[EQ_VX + a] + f.cof(a, i) * qL[EQ_VX + i];
            qRh[EQ_VX + a] = qRh[EQ_VX + a] + f.cof(a, i) * qR[EQ_VX + i];
        }
    }
    if (!srhd_local(g1, gam, qLh, uL, fL, lLmin, lLmax)) return false;
    if (!srhd_local(g1, gam, qRh, uR, fR, lRmin, lRmax)) return false;
    const R ll = rmin(lLmin, lRmin), lr = rmax(lLmax, lRmax);
    const R w = f.wface();
    if (!finite(w) || lr - ll <= (R)1.0e-13) return false;
    if (w <= ll) return local_to_coord<DIR>(uL, fL, w, f, F);
    if (w >= lr) return local_to_coord<DIR>(uR, fR, w, f, F);
    R uh[NEQ], fh[NEQ];
    R den = lr - ll;
    const R rhll = rcp(den);
    for (int e = 0; e < NEQ; e++) {
        uh[e] = quot(lr * uR[e] - ll * uL[e] + fL[e] - fR[e], den, rhll);
        fh[e] = quot(lr * fL[e] - ll * fR[e] + lr * ll * (uR[e] - uL[e]), den, rhll);
    }
    for (int e = 0; e < NEQ; e++)
        if (!finite(uh[e]) || !finite(fh[e])) return false;
    const R Eh = uh[EQ_PR] + uh[EQ_DE], FEh = fh[EQ_PR] + fh[EQ_DE];
    const R A = FEh, B = -(Eh + fh[EQ_VX]), Cq = uh[EQ_VX];
    R disc = B * B - (R)4 * A * Cq;
    const R dscale = rmax(rmax(Lim<R>::tiny(), B * B), fabs((R)4 * A * Cq));
    if (!finite(disc) || disc < (R)-1.0e-12 * dscale) return false;
    disc = rmax((R)0, disc);
    const R cscale = rmax(rmax(Lim<R>::tiny(), fabs(B)), fabs(Cq));
    const bool quad = fabs(A) > (R)1.0e-13 * cscale;
    R ls;
    if (!quad) {
        if (fabs(B) <= (R)1.0e-14 * rmax(Lim<R>::tiny(), fabs(Cq))) return false;
        ls = rdiv(-Cq, B);
    } else {
        den = -B - sqrtr(disc);
        if (fabs(den) > (R)1.0e-14 * rmax(rmax(Lim<R>::tiny(), fabs(B)), sqrtr(disc))) {
            ls = rdiv(den, (R)2 * A);
        } else {
            den = -B + sqrtr(disc);
            if (fabs(den) <= (R)1.0e-14 * rmax(rmax(Lim<R>::tiny(), fabs(B)), sqrtr(disc)))
                return false;
            ls = rdiv((R)2 * Cq, den);
        }
    }
    const R rtol = (R)1.0e-10 * rmax(rmax((R)1, fabs(ll)), fabs(lr));
    bool valid = finite(ls) && ls >= ll - rtol && ls <= lr + rtol && fabs(ls) <= (R)1 + rtol;
    if (!valid && quad) {
        const R ra = rdiv(-B + sqrtr(disc), (R)2 * A);
        if (finite(ra) && ra >= ll - rtol && ra <= lr + rtol && fabs(ra) <= (R)1 + rtol) {
            ls = ra;
            valid = true;
        }
    }
    if (!valid) return false;
    ls = rmin(lr, rmax(ll, ls));
    const R AL = ll * (uL[EQ_PR] + uL[EQ_DE]) - uL[EQ_VX];
    const R BL = uL[EQ_VX] * (ll - qLh[EQ_VX]) - qLh[EQ_PR];
    const R AR = lr * (uR[EQ_PR] + uR[EQ_DE]) - uR[EQ_VX];
    const R BR = uR[EQ_VX] * (lr - qRh[EQ_VX]) - qRh[EQ_PR];
    den = (R)1 - ll * ls;
    if (fabs(den) <= (R)1.0e-13) return false;
    const R psL = rdiv(AL * ls - BL, den);
    den = (R)1 - lr * ls;
    if (fabs(den) <= (R)1.0e-13) return false;
    const R psR = rdiv(AR * ls - BR, den);
    const R sscale = rmax(rmax(rmax(rmax(Lim<R>::tiny(), fabs(uL[EQ_PR] + uL[EQ_DE])),
                                    fabs(uR[EQ_PR] + uR[EQ_DE])),
                               fabs(uL[EQ_VX])),
                          fabs(uR[EQ_VX]));
    const R ptol = (R)1.0e-8 * rmax(rmax(rmax(rmax(Lim<R>::tiny(), fabs(qLh[EQ_PR])),
                                              fabs(qRh[EQ_PR])),
                                         fabs(psL)),
                                    fabs(psR)) +
                   (R)1.0e2 * Lim<R>::eps() * sscale;
    if (!finite(psL) || !finite(psR) || fabs(psL - psR) > ptol) return false;
    const R ps = (R)0.5 * (psL + psR);
    if (!finite(ps) || fabs(ll - ls) <= (R)1.0e-13 || fabs(lr - ls) <= (R)1.0e-13) return false;
    const R pv[NEQ] = {0, ls, 1, 0, 0};
    R usL[NEQ], usR[NEQ], fsL[NEQ], fsR[NEQ];
    const R dsL = ll - ls, dsR = lr - ls, rdsL = rcp(dsL), rdsR = rcp(dsR);
    for (int e = 0; e < NEQ; e++) {
        if (GRHD_CZERO(pv[e])) {
            usL[e] = quot(ll * uL[e] - fL[e], dsL, rdsL);
            usR[e] = quot(lr * uR[e] - fR[e], dsR, rdsR);
        } else {
            usL[e] = quot(ll * uL[e] - fL[e] + ps * pv[e], dsL, rdsL);
            usR[e] = quot(lr * uR[e] - fR[e] + ps * pv[e], dsR, rdsR);
        }
    }
    for (int e = 0; e < NEQ; e++) {
        fsL[e] = fL[e] + ll * (usL[e] - uL[e]);
        fsR[e] = fR[e] + lr * (usR[e] - uR[e]);
    }
    if (!(admissible(usL) & admissible(usR))) return false;
    for (int e = 0; e < NEQ; e++)
        if (!finite(fsL[e]) || !finite(fsR[e])) return false;
    if (w <= ls) return local_to_coord<DIR>(usL, fsL, w, f, F);
    return local_to_coord<DIR>(usR, fsR, w, f, F);
}

// resolve_riemann_problem (fluxes.f90:56-72).
template <typename R, int DIR, bool SRHD, typename FG>
GRHD_HD void riemann(int rs, R g1, R gam, const R qL[NEQ], const R qR[NEQ], const FG &f,
                     R F[NEQ])
{
    if (rs == 2) {
        if (!(DIR >= 1 && DIR <= 3)) {
            hlle<R, DIR, SRHD>(g1, gam, qL, qR, f, F);
            return;
        }
        if (hllc_try<R, DIR>(g1, gam, qL, qR, f, F)) return;
    }
    hlle<R, DIR, SRHD>(g1, gam, qL, qR, f, F);
}

// -------------------------------------------------------- reconstruction --
template <typename R> GRHD_HD R minmod2(R a, R b)
{   // reconstruction.f90:28-32
    return (R)0.5 * (sgn1(a) + sgn1(b)) * rmin(fabs(a), fabs(b));
}
template <typename R> GRHD_HD R minmod3(R a, R b, R c)
{   // reconstruction.f90:34-38
    return sgn1(a) * rmax((R)0, rmin(rmin(fabs(a), sgn1(a) * b), sgn1(a) * c));
}
template <typename R> GRHD_HD R minmod4(R a, R b, R c, R d)
{   // reconstruction.f90:40-45
    return (R)0.125 * (sgn1(a) + sgn1(b)) * fabs(sgn1(a) + sgn1(c)) * fabs(sgn1(a) + sgn1(d)) *
           rmin(rmin(rmin(fabs(a), fabs(b)), fabs(c)), fabs(d));
}
// weno3_interface with the "dynamic epsilon" (reconstruction.f90:141-181).
template <typename R> GRHD_HD R weno3(R qm1, R q0, R qp1)
{
    const R d0 = (R)(2.0 / 3.0), d1 = (R)(1.0 / 3.0), eps_min = (R)1.0e-14;
    const R p0 = (R)0.5 * q0 + (R)0.5 * qp1;
    const R p1 = (R)-0.5 * qm1 + (R)1.5 * q0;
    const R b0 = (q0 - qp1) * (q0 - qp1);
    const R b1 = (qm1 - q0) * (qm1 - q0);
    const R sens = rdiv(fabs(b0 - b1), rmax(b0, b1) + eps_min);
    const R epsh = eps_min + (qm1 * qm1 + q0 * q0 + qp1 * qp1) * exp((R)-10 * sens);
    const R a0 = rdiv(d0, (epsh + b0) * (epsh + b0));
    const R a1 = rdiv(d1, (epsh + b1) * (epsh + b1));
    const R as = a0 + a1, ras = rcp(as);
    const R w0 = quot(a0, as, ras), w1 = quot(a1, as, ras);
    return w0 * p0 + w1 * p1;
}
// weno5_interface, WENO5-Z (reconstruction.f90:187-215).
template <typename R> GRHD_HD R weno5(R qm2, R qm1, R q0, R qp1, R qp2)
{
    const R eps = (R)1.0e-14, d0 = (R)0.3, d1 = (R)0.6, d2 = (R)0.1;
    const R t0 = q0 - (R)2 * qp1 + qp2, u0 = (R)3 * q0 - (R)4 * qp1 + qp2;
    const R t1 = qm1 - (R)2 * q0 + qp1, u1 = qm1 - qp1;
    const R t2 = qm2 - (R)2 * qm1 + q0, u2 = qm2 - (R)4 * qm1 + (R)3 * q0;
    const R c13 = (R)(13.0 / 12.0);
    const R b0 = c13 * (t0 * t0) + (R)0.25 * (u0 * u0);
    const R b1 = c13 * (t1 * t1) + (R)0.25 * (u1 * u1);
    const R b2 = c13 * (t2 * t2) + (R)0.25 * (u2 * u2);
    const R tau5 = fabs(b0 - b2);
    const R a0 = d0 * ((R)1 + rdiv(tau5, eps + b0));
    const R a1 = d1 * ((R)1 + rdiv(tau5, eps + b1));
    const R a2 = d2 * ((R)1 + rdiv(tau5, eps + b2));
    const R as = a0 + a1 + a2, ras = rcp(as);
    const R w0 = quot(a0, as, ras), w1 = quot(a1, as, ras), w2 = quot(a2, as, ras);
    const R p0 = (R)(1.0 / 3.0) * q0 + (R)(5.0 / 6.0) * qp1 - (R)(1.0 / 6.0) * qp2;
    const R p1 = -(R)(1.0 / 6.0) * qm1 + (R)(5.0 / 6.0) * q0 + (R)(1.0 / 3.0) * qp1;
    const R p2 = (R)(1.0 / 3.0) * qm2 - (R)(7.0 / 6.0) * qm1 + (R)(11.0 / 6.0) * q0;
    return w0 * p0 + w1 * p1 + w2 * p2;
}
// mp5_interface, Suresh-Huynh MP5 with alpha = 4 (reconstruction.f90:222-254).
template <typename R> GRHD_HD R mp5(R qm2, R qm1, R q0, R qp1, R qp2)
{
    const R alpha = (R)4, eps = (R)1.0e-10;
    const R vL = quot((R)2 * qm2 - (R)13 * qm1 + (R)47 * q0 + (R)27 * qp1 - (R)3 * qp2, (R)60,
                      (R)(1.0 / 60.0));
    const R vmp = q0 + minmod2(qp1 - q0, alpha * (q0 - qm1));
    if ((vL - q0) * (vL - vmp) < eps) return vL;
    const R djm1 = qm2 - (R)2 * qm1 + q0;
    const R dj = qm1 - (R)2 * q0 + qp1;
    const R djp1 = q0 - (R)2 * qp1 + qp2;
    const R dm4p = minmod4((R)4 * dj - djp1, (R)4 * djp1 - dj, dj, djp1);
    const R dm4m = minmod4((R)4 * dj - djm1, (R)4 * djm1 - dj, dj, djm1);
    const R vul = q0 + alpha * (q0 - qm1);
    const R vav = (R)0.5 * (q0 + qp1);
    const R vmd = vav - (R)0.5 * dm4p;
    const R vlc = q0 + (R)0.5 * (q0 - qm1) + quot((R)4 * dm4m, (R)3, (R)(1.0 / 3.0));
    const R vmin = rmax(rmin(rmin(q0, qp1), vmd), rmin(rmin(q0, vul), vlc));
    const R vmax = rmin(rmax(rmax(q0, qp1), vmd), rmax(rmax(q0, vul), vlc));
    return vL + minmod2(vmin - vL, vmax - vL);
}
// tvd_slope (reconstruction.f90:259-278).
template <typename R> GRHD_HD R tvd_slope(R a, R b, int lim)
{
    if (lim == 2) {
        if (a * b <= (R)0) return 0;
        return sgn1(a) * rmax(rmin((R)2 * fabs(a), fabs(b)), rmin(fabs(a), (R)2 * fabs(b)));
    }
    if (lim == 3) return minmod3((R)2 * b, (R)0.5 * (a + b), (R)2 * a);
    return minmod2(a, b);
}

// Face schemes compiled as constants: scheme 0 reads the reconstruction, the
// shock sensor and the Riemann solver from Params at run time (every
// combination); the others fix them, which removes the untaken branches and
// their registers from the face kernels of the configurations run most:
// 1 = MP5 + shock sensor + HLLC, 2 = WENO5 + HLLE without the sensor.
struct Scheme {
    int rec, sensor, rs;
};
constexpr Scheme kSchemes[] = {{0, 0, 0}, {4, 1, 2}, {5, 0, 1}};
// ghost depth of a reconstruction (initialization.f90:27-52)
GRHD_HD constexpr int ghost_depth(int rec) { return rec == 1 ? 1 : (rec == 2 || rec == 3) ? 2 : 3; }

// The shock sensor's choice for one face (reconstruction.f90:292-310, 331-520): the
// relative pressure and density jumps over the five cells around the face
// select Godunov, TVD-MC, WENO3 or the base scheme.  s[v][o + 2] holds the
// primitive v at offset o = -2..3 from the face's left cell; i is the face
// index (0..nodes) along direction dir (1..3).
struct SensorPick {
    bool god, tvdmc, w3, mp5w3;
};
template <typename R>
GRHD_HD SensorPick sensor_pick(const Params &P, const R s[NEQ][6], int i, int nodes, int dir)
{
#define S_(v, o) s[v][(o) + 2]
    const bool mink = (P.mk <= MK_MINK_SPH_LOG);
    const R rf = (R)P.rho_floor, pf = (R)P.p_floor;
    const bool radial0 = (!mink && dir == 1 && i <= 0);
    R pmax = S_(EQ_PR, -2), pmin = S_(EQ_PR, -2), rmx = S_(EQ_DE, -2), rmn = S_(EQ_DE, -2);
    for (int o = -1; o <= 2; o++) {
        pmax = rmax(pmax, S_(EQ_PR, o));
        pmin = rmin(pmin, S_(EQ_PR, o));
        rmx = rmax(rmx, S_(EQ_DE, o));
        rmn = rmin(rmn, S_(EQ_DE, o));
    }
    const R pj = (pmax > (R)1.0e3 * pf) ? rdiv(pmax - pmin, rmax(pmin, pf)) : (R)0;
    const R rj = (rmx > (R)1.0e3 * rf) ? rdiv(rmx - rmn, rmax(rmn, rf)) : (R)0;
    SensorPick sp;
    sp.god = radial0 || S_(EQ_DE, 0) <= (R)1.0e2 * rf || S_(EQ_DE, 1) <= (R)1.0e2 * rf || pj >= (R)10 ||
             rj >= (R)10;
    sp.tvdmc = (pj >= (R)0.80 || rj >= (R)0.50);
    sp.w3 = (pj >= (R)0.40 || rj >= (R)0.15);
    sp.mp5w3 = sp.w3 || (!mink && dir == 1 && i <= 3) || (!mink && dir == 2 && (i <= 2 || i >= nodes - 2));
    return sp;
#undef S_
}

// One face of reconstruct_1d_core (reconstruction.f90:285-657), with the
// stencil s and face index i of sensor_pick.
template <typename R, int SCH = 0>
GRHD_HD void recon_face(const Params &P, const R s[NEQ][6], int i, int nodes, int dir,
                        R qL[NEQ], R qR[NEQ])
{
#define S_(v, o) s[v][(o) + 2]
    const int rec = SCH ? kSchemes[SCH].rec : P.rec;
    const bool sensor = SCH ? (kSchemes[SCH].sensor != 0) : (P.sensor != 0);
    const bool mink = (P.mk <= MK_MINK_SPH_LOG);
    const R rf = (R)P.rho_floor, pf = (R)P.p_floor;
    const bool radial0 = (!mink && dir == 1 && i <= 0);
    if (rec == 1) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
            qL[v] = S_(v, 0);
            qR[v] = S_(v, 1);
        }
    } else if (sensor) {
        const SensorPick sp = sensor_pick<R>(P, s, i, nodes, dir);
        const bool god = sp.god, tvdmc = sp.tvdmc, w3 = sp.w3, mp5w3 = sp.mp5w3;
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
            if (god) {
                qL[v] = S_(v, 0);
                qR[v] = S_(v, 1);
            } else if (rec == 2 || tvdmc) {
                const int lim = (rec == 2) ? P.lim : 3;
                qL[v] = S_(v, 0) + (R)0.5 * tvd_slope(S_(v, 0) - S_(v, -1), S_(v, 1) - S_(v, 0), lim);
                qR[v] = S_(v, 1) - (R)0.5 * tvd_slope(S_(v, 1) - S_(v, 0), S_(v, 2) - S_(v, 1), lim);
            } else if (rec == 3 || (rec == 4 && mp5w3) || (rec == 5 && w3)) {
                qL[v] = weno3(S_(v, -1), S_(v, 0), S_(v, 1));
                qR[v] = weno3(S_(v, 2), S_(v, 1), S_(v, 0));
            } else if (rec == 4) {
                qL[v] = mp5(S_(v, -2), S_(v, -1), S_(v, 0), S_(v, 1), S_(v, 2));
                qR[v] = mp5(S_(v, 3), S_(v, 2), S_(v, 1), S_(v, 0), S_(v, -1));
            } else {
                qL[v] = weno5(S_(v, -2), S_(v, -1), S_(v, 0), S_(v, 1), S_(v, 2));
                qR[v] = weno5(S_(v, 3), S_(v, 2), S_(v, 1), S_(v, 0), S_(v, -1));
            }
        }
    } else {
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
            if (rec == 2) {
                qL[v] = S_(v, 0) + (R)0.5 * tvd_slope(S_(v, 0) - S_(v, -1), S_(v, 1) - S_(v, 0), P.lim);
                qR[v] = S_(v, 1) - (R)0.5 * tvd_slope(S_(v, 1) - S_(v, 0), S_(v, 2) - S_(v, 1), P.lim);
            } else if (radial0) {
                qL[v] = S_(v, 0);
                qR[v] = S_(v, 1);
            } else if (rec == 3) {
                qL[v] = weno3(S_(v, -1), S_(v, 0), S_(v, 1));
                qR[v] = weno3(S_(v, 2), S_(v, 1), S_(v, 0));
            } else if (rec == 4) {
                qL[v] = mp5(S_(v, -2), S_(v, -1), S_(v, 0), S_(v, 1), S_(v, 2));
                qR[v] = mp5(S_(v, 3), S_(v, 2), S_(v, 1), S_(v, 0), S_(v, -1));
            } else {
                qL[v] = weno5(S_(v, -2), S_(v, -1), S_(v, 0), S_(v, 1), S_(v, 2));
                qR[v] = weno5(S_(v, 3), S_(v, 2), S_(v, 1), S_(v, 0), S_(v, -1));
            }
        }
    }
    // positivity floors (reconstruction.f90:647-655)
    qL[EQ_DE] = rmax(qL[EQ_DE], rf);
    qR[EQ_DE] = rmax(qR[EQ_DE], rf);
    qL[EQ_PR] = rmax(qL[EQ_PR], pf);
    qR[EQ_PR] = rmax(qR[EQ_PR], pf);
#undef S_
}

// -------------------------------------------------------------- con2prim --
// eval_V_and_dV (equations.f90:212-220) and eval_residual (223-246).
GRHD_HD void eval_V(double g1, double W, double D, double tau, double &V, double &dV)
{
    const double den = g1 * (W * W) - 1.0;
    V = div64(g1 * (W * W) * (tau + D) - D * W, den);
    dV = div64(g1 * W * (D * W - 2.0 * (tau + D)) + D, den * den);
}
GRHD_HD void eval_res(double g1, double W, double D, double tau, double Q, double &f, double &df)
{
    double V, dV;
    eval_V(g1, W, D, tau, V, dV);
    f = (V * V) * (1.0 - rcp64(W * W)) - Q;
    df = div64(2.0 * V * (dV * (W * W * W - W) + V), W * W * W);
}

// root_brent (equations.f90:343-424).
GRHD_HD double root_brent(double g1, double rho_floor, double D, double tau, double Q)
{
    const double x1 = 1.0, xacc = 1.0e-8, EPS = 3.0e-8;
    const double wg = (sqrtr(Q) + D) / rmax(D, rho_floor);
    const double x2 = rmax(1.0e4, wg * 2.0);
    double a, b, cc, d = 0.0, e = 0.0, fa, fb, fc, df;
    eval_res(g1, x1, D, tau, Q, fa, df);
    eval_res(g1, x2, D, tau, Q, fb, df);
    if (sgn1(fa) * sgn1(fb) > 0.0) return -1.0;
    a = x1;
    b = x2;
    cc = b;
    fc = fb;
    for (int it = 1; it <= 100; it++) {
        if (sgn1(fb) * sgn1(fc) > 0.0) {
            cc = a;
            fc = fa;
            d = b - a;
            e = d;
        }
        if (fabs(fc) < fabs(fb)) {
            a = b;
            b = cc;
            cc = a;
            fa = fb;
            fb = fc;
            fc = fa;
        }
        const double tol1 = 2.0 * EPS * fabs(b) + 0.5 * xacc;
        const double xm = 0.5 * (cc - b);
        if (fabs(xm) <= tol1 || fabs(fb) < 1.0e-14) return b;
        if (fabs(e) >= tol1 && fabs(fa) > fabs(fb)) {
            double o, q, r;
            const double y1 = fb / fa;
            if (fabs(a - cc) < 1.0e-14) {
                o = 2.0 * xm * y1;
                q = 1.0 - y1;
            } else {
                q = fa / fc;
                r = fb / fc;
                o = y1 * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0));
                q = (q - 1.0) * (r - 1.0) * (y1 - 1.0);
            }
            if (o > 0.0) q = -q;
            o = fabs(o);
            if (2.0 * o < rmin(3.0 * xm * q - fabs(tol1 * q), fabs(e * q))) {
                e = d;
                d = o / q;
            } else {
                d = xm;
                e = d;
            }
        } else {
            d = xm;
            e = d;
        }
        a = b;
        fa = fb;
        if (fabs(d) > tol1) b = b + d;
        else b = b + copysign(tol1, xm);
        eval_res(g1, b, D, tau, Q, fb, df);
    }
    return -1.0;
}

// root_safe_newton (equations.f90:249-340).  *brent is set when the
// Brent fallback runs.  w_start > 0 starts the safeguarded iteration there
// instead of at the bracket midpoint (warm start from the previous Lorentz
// factor, used by the mixed-precision path only; the bracket, the
// Newton/bisection safeguard, the tolerances and the Brent fallback are
// unchanged).
GRHD_HD double root_newton(double g1, double rho_floor, double D, double tau, double Q, int &brent,
                           double w_start = 0.0)
{
    const double x1 = 1.0, xacc = 1.0e-10;
    const double wg = (sqrtr(Q) + D) / rmax(D, rho_floor);
    const double x2 = rmax(1.0e3, wg * 2.0);
    double xl, xh, f, df, flo, fhi, dx, dxold, root, tmp;
    // The residual at the lower bracket W = 1: eval_res's operations with
    // W = 1 (multiplications by 1 are exact, 1 - 1/1^2 = 0), which leaves
    // f(1) = V(1)^2 0 - Q, the same value in every case (NaN when V(1) or its
    // square is not finite), without dV and the derivative, which the solve
    // overwrites before use.
    const double V1 = div64(g1 * (tau + D) - D, g1 - 1.0);
    flo = (V1 * V1) * 0.0 - Q;
    eval_res(g1, x2, D, tau, Q, fhi, df);
    if (fabs(flo) < 1.0e-14) return x1;
    if (fabs(fhi) < 1.0e-14) return x2;
    if (sgn1(flo) * sgn1(fhi) > 0.0) return -1.0;
    if (flo < 0.0) {
        xl = x1;
        xh = x2;
    } else {
        xh = x1;
        xl = x2;
    }
    root = (w_start > x1 && w_start < x2) ? w_start : 0.5 * (x1 + x2);
    dxold = fabs(x2 - x1);
    dx = dxold;
    eval_res(g1, root, D, tau, Q, f, df);
    for (int it = 1; it <= 100; it++) {
        if (((root - xh) * df - f) * ((root - xl) * df - f) > 0.0 || fabs(2.0 * f) > fabs(dxold * df)) {
            dxold = dx;
            dx = 0.5 * (xh - xl);
            root = xl + dx;
            if (fabs(xl - root) < 1.0e-14) return root;
        } else {
            dxold = dx;
            dx = div64(f, df);
            tmp = root;
            root = root - dx;
            if (fabs(tmp - root) < 1.0e-14) return root;
        }
        if (fabs(dx) < xacc) return root;
        eval_res(g1, root, D, tau, Q, f, df);
        if (f < 0.0) xl = root;
        else xh = root;
    }
    brent = 1;
    return root_brent(g1, rho_floor, D, tau, Q);
}

// r[0] x + r[1] y + r[2] z, left to right, without the terms of structural
// zeros of r.
GRHD_HD double dot3(const double r[3], double x, double y, double z)
{
    double s = r[0] * x;
    if (!GRHD_CZERO(r[1])) s = GRHD_CZERO(r[0]) ? r[1] * y : s + r[1] * y;
    if (!GRHD_CZERO(r[2])) s = (GRHD_CZERO(r[0]) && GRHD_CZERO(r[1])) ? r[2] * z : s + r[2] * z;
    return s;
}

// cons_to_prim (equations.f90:94-205).  c is repaired in place; returns true
// when the cell was reset to the atmosphere.
GRHD_HD bool c2p(double g1, double rho_floor, double p_floor, const double g[3][3],
                 const double gi[3][3], double sqg, double c[NEQ], double q[NEQ], int &brent,
                 double w_start = 0.0, bool *repaired = nullptr)
{
    const double vmax = 1.0 - 1.0e-5;
    const double Df = rho_floor * sqg, tf = p_floor * (g1 - 1.0) * sqg;
    bool atm = (c[EQ_DE] <= Df);
    if (!atm) {
        const double S1 = c[EQ_VX], S2 = c[EQ_VY], S3 = c[EQ_VZ];
        const double Ssq = S1 * dot3(gi[0], S1, S2, S3) + S2 * dot3(gi[1], S1, S2, S3) +
                           S3 * dot3(gi[2], S1, S2, S3);
        const double E = c[EQ_PR] + c[EQ_DE];
        const double Smax = rmax(0.0, E * E - c[EQ_DE] * c[EQ_DE]) * vmax;
        if (Ssq > Smax && Ssq > 1.0e-20) {
            const double sf = sqrtr(Smax / Ssq);
            if (repaired) *repaired = true;
            c[EQ_VX] = c[EQ_VX] * sf;
            c[EQ_VY] = c[EQ_VY] * sf;
            c[EQ_VZ] = c[EQ_VZ] * sf;
        }
        double Q = 0.0;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                if (!GRHD_CZERO(gi[a][b])) Q = Q + gi[a][b] * c[EQ_VX + a] * c[EQ_VX + b];
        const double W = root_newton(g1, rho_floor, c[EQ_DE], c[EQ_PR], Q, brent, w_start);
        if (W < 1.0) {
            atm = true;
        } else {
            double V, dV;
            eval_V(g1, W, c[EQ_DE], c[EQ_PR], V, dV);
            if (V <= 0.0 || isnan(V)) {
                atm = true;
            } else {
                q[EQ_DE] = div64(c[EQ_DE], W * sqg);
                q[EQ_PR] = div64(V - c[EQ_PR] - c[EQ_DE], sqg);
                const double rV = rcp64(V);
                q[EQ_VX] = quot(dot3(gi[0], c[EQ_VX], c[EQ_VY], c[EQ_VZ]), V, rV);
                q[EQ_VY] = quot(dot3(gi[1], c[EQ_VX], c[EQ_VY], c[EQ_VZ]), V, rV);
                q[EQ_VZ] = quot(dot3(gi[2], c[EQ_VX], c[EQ_VY], c[EQ_VZ]), V, rV);
                double vf = 0.0;
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                        if (!GRHD_CZERO(g[a][b])) vf = vf + g[a][b] * q[EQ_VX + a] * q[EQ_VX + b];
                atm = (q[EQ_DE] < rho_floor || q[EQ_PR] < p_floor || isnan(q[EQ_DE]) ||
                       isnan(q[EQ_PR]) || isnan(q[EQ_VX]) || isnan(q[EQ_VY]) || isnan(q[EQ_VZ]) ||
                       isnan(vf) || vf >= vmax);
            }
        }
    }
    if (atm) {
        if (repaired) *repaired = true;
        q[EQ_DE] = rho_floor;
        q[EQ_PR] = p_floor;
        q[EQ_VX] = q[EQ_VY] = q[EQ_VZ] = 0.0;
        c[EQ_DE] = Df;
        c[EQ_PR] = tf;
        c[EQ_VX] = c[EQ_VY] = c[EQ_VZ] = 0.0;
    }
    return atm;
}

// ---------------------------------------------------------------- sources --
// calc_sources (equations.f90:509-552) with T^{mu nu} from
// calculate_stress_energy_tensor (metrics.f90:1251-1310).  The d_phi
// derivatives vanish for every metric here, so S_3 = 0.
GRHD_HD void sources(double g1, const CentreMetric &m, const double q[NEQ], double src[NEQ])
{
    double v[3] = {q[EQ_VX], q[EQ_VY], q[EQ_VZ]}, u4[4], T[4][4];
    const double vsq = vsq_clamp<double>(Mat3<double>{m.g}, v);
    const double h = 1.0 + g1 * q[EQ_PR] / q[EQ_DE];
    const double W = 1.0 / sqrtr(1.0 - vsq);
    const double ra = 1.0 / m.alpha;
    u4[0] = quot(W, m.alpha, ra);
    for (int i = 0; i < 3; i++)
        u4[i + 1] = GRHD_CZERO(m.beta[i]) ? W * v[i] : W * (v[i] - quot(m.beta[i], m.alpha, ra));
    for (int mu = 0; mu < 4; mu++)
        for (int nu = mu; nu < 4; nu++) {
            T[mu][nu] = q[EQ_DE] * h * u4[mu] * u4[nu];
            if (!GRHD_CZERO(m.gm[mu][nu])) T[mu][nu] = T[mu][nu] + (q[EQ_PR] * m.gm[mu][nu]);
            T[nu][mu] = T[mu][nu];
        }
    for (int e = 0; e < NEQ; e++) src[e] = 0.0;
    for (int d = 0; d < 2; d++) {
        double s = 0.0;
        for (int mu = 0; mu < 4; mu++)
            for (int nu = 0; nu < 4; nu++)
                if (!GRHD_CZERO(m.dg[mu][nu][d])) s = s + 0.5 * T[mu][nu] * m.dg[mu][nu][d];
        src[EQ_VX + d] = s;
    }
    double st = 0.0;
    for (int mu = 0; mu < 4; mu++) {
        if (!GRHD_CZERO(m.dlna[mu])) st = st + T[mu][0] * m.dlna[mu];
        for (int nu = 0; nu < 4; nu++)
            if (!GRHD_CZERO(m.ch0[mu][nu])) st = st - T[mu][nu] * m.ch0[mu][nu];
    }
    src[EQ_PR] = st;
    for (int e = 0; e < NEQ; e++)
        if (!GRHD_CZERO(src[e])) src[e] = src[e] * m.alpha * m.sqg;
    src[EQ_PR] = src[EQ_PR] * m.alpha;
}

}  // namespace grhd
