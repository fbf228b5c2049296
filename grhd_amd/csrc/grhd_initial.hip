// grhd_initial.hip -- initial data on the host (conditions.f90:26-214, 229-582).
//
// Runs once per problem, so it is plain scalar C++; it shares the analytic
// metric with the kernels (grhd_physics.hpp, host-callable).  The result is a
// full ghosted primitive array in the reference layout.
#include "grhd_initial.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace grhd {

namespace {

struct HostGrid {
    const Params &P;
    int ng, NXg, NYg, NZg;
    explicit HostGrid(const Params &p)
        : P(p), ng(p.ng), NXg(p.nx + 2 * p.ng + 1), NYg(p.ny + 2 * p.ng + 1), NZg(p.nz + 2 * p.ng + 1) {}
    size_t at(int v, int i, int j, int k) const
    {
        return ((((size_t)(k + ng) * NYg + (size_t)(j + ng)) * NXg + (size_t)(i + ng)) * NEQ) + v;
    }
};

template <int MK> void face_metric(const Params &P, double x, double y, FaceMetric &m)
{
    metric_face<MK>(P, coord_of(MK, x, y), m);
}

void metric_any(const Params &P, double x, double y, FaceMetric &m)
{
    switch (P.mk) {
    case MK_MINK_CART: face_metric<MK_MINK_CART>(P, x, y, m); break;
    case MK_MINK_CYL: face_metric<MK_MINK_CYL>(P, x, y, m); break;
    case MK_MINK_SPH: face_metric<MK_MINK_SPH>(P, x, y, m); break;
    case MK_MINK_SPH_LOG: face_metric<MK_MINK_SPH_LOG>(P, x, y, m); break;
    case MK_EF: face_metric<MK_EF>(P, x, y, m); break;
    case MK_EF_LOG: face_metric<MK_EF_LOG>(P, x, y, m); break;
    case MK_KS: face_metric<MK_KS>(P, x, y, m); break;
    default: face_metric<MK_KS_LOG>(P, x, y, m); break;
    }
}

bool log_radius(const Params &P)
{
    return P.mk == MK_EF_LOG || P.mk == MK_KS_LOG || P.mk == MK_MINK_SPH_LOG;
}

// fishbone_metric_components (conditions.f90:575-582).
void fm_components(const FaceMetric &m, double &g_tt, double &g_tphi)
{
    double gb[3];
    for (int a = 0; a < 3; a++) {
        gb[a] = 0.0;
        for (int b = 0; b < 3; b++) gb[a] = gb[a] + m.g[a][b] * m.beta[b];
    }
    g_tt = -(m.alpha * m.alpha) + (m.beta[0] * gb[0] + m.beta[1] * gb[1] + m.beta[2] * gb[2]);
    g_tphi = m.g[2][0] * m.beta[0] + m.g[2][1] * m.beta[1] + m.g[2][2] * m.beta[2];
}

}  // namespace

// prepare_fishbone_moncrief (conditions.f90:470-510).
bool fishbone_constants(const Params &P, double r_in, double r_max_dens, double &l_ang, double &W_in)
{
    const double M = P.mass, a = P.spin;
    l_ang = 0.0;
    W_in = 0.0;
    if (M <= 0.0 || r_in <= 0.0 || r_max_dens <= r_in) return false;
    const double sq = std::sqrt(M * r_max_dens);
    const double num = sq * (r_max_dens * r_max_dens - 2.0 * a * sq + a * a);
    const double den = r_max_dens * r_max_dens - 2.0 * M * r_max_dens + a * sq;
    if (!std::isfinite(den) || std::fabs(den) <= DBL_MIN) return false;
    l_ang = num / den;
    if (!std::isfinite(l_ang)) return false;
    const double coord = log_radius(P) ? std::log(r_in) : r_in;
    FaceMetric m;
    metric_any(P, coord, M_PI / 2.0, m);
    double g_tt, g_tphi;
    fm_components(m, g_tt, g_tphi);
    const double nu = g_tphi * g_tphi - g_tt * m.g[2][2];
    const double du = m.g[2][2] + 2.0 * l_ang * g_tphi + l_ang * l_ang * g_tt;
    if (nu <= 0.0 || du <= 0.0) return false;
    const double ut2 = nu / du;
    if (!std::isfinite(ut2) || ut2 <= 0.0) return false;
    W_in = 0.5 * std::log(ut2);
    return std::isfinite(W_in);
}

// evaluate_fishbone_moncrief_state (conditions.f90:516-572).
static void fishbone_state(const Params &P, double r_loc, double r_in, double K, double l_ang,
                           double W_in, const FaceMetric &m, double *q)
{
    for (int v = 0; v < NEQ; v++) q[v] = 0.0;
    q[EQ_DE] = P.rho_floor;
    q[EQ_PR] = P.p_floor;
    if (r_loc < r_in || K <= 0.0 || m.alpha <= 0.0) return;
    double g_tt, g_tphi;
    fm_components(m, g_tt, g_tphi);
    const double nu = g_tphi * g_tphi - g_tt * m.g[2][2];
    const double du = m.g[2][2] + 2.0 * l_ang * g_tphi + l_ang * l_ang * g_tt;
    if (nu <= 0.0 || du <= 0.0) return;
    const double ut2 = nu / du;
    if (!std::isfinite(ut2) || ut2 <= 0.0) return;
    const double W_r = 0.5 * std::log(ut2);
    const double h = std::exp(W_in - W_r);
    if (!std::isfinite(h) || h <= 1.0) return;
    const double gam = P.gam;
    const double rho = std::pow((h - 1.0) * (gam - 1.0) / (K * gam), 1.0 / (gam - 1.0));
    const double pr = K * std::pow(rho, gam);
    if (!std::isfinite(rho) || !std::isfinite(pr) || rho <= P.rho_floor || pr <= P.p_floor) return;
    const double od = m.g[2][2] + l_ang * g_tphi;
    if (!std::isfinite(od) || std::fabs(od) <= DBL_MIN) return;
    const double Om = -(g_tphi + l_ang * g_tt) / od;
    if (!std::isfinite(Om)) return;
    double c[NEQ] = {rho, pr, m.beta[0] / m.alpha, m.beta[1] / m.alpha, (Om + m.beta[2]) / m.alpha};
    for (int v = 0; v < NEQ; v++)
        if (!std::isfinite(c[v])) return;
    for (int v = 0; v < NEQ; v++) q[v] = c[v];
}

double fishbone_K_for_unit_density(const Params &P, double r_in, double r_c)
{
    // rho(r_c) = ((h_c - 1)(G - 1)/(K G))^(1/(G-1)) = 1  =>  K = (h_c - 1)(G - 1)/G with
    // h_c = exp(W_in - W(r_c, pi/2)) (conditions.f90:540-546).
    double l_ang, W_in;
    if (!fishbone_constants(P, r_in, r_c, l_ang, W_in)) return -1.0;
    const double coord = log_radius(P) ? std::log(r_c) : r_c;
    FaceMetric m;
    metric_any(P, coord, M_PI / 2.0, m);
    double g_tt, g_tphi;
    fm_components(m, g_tt, g_tphi);
    const double nu = g_tphi * g_tphi - g_tt * m.g[2][2];
    const double du = m.g[2][2] + 2.0 * l_ang * g_tphi + l_ang * l_ang * g_tt;
    const double h = std::exp(W_in - 0.5 * std::log(nu / du));
    return (h - 1.0) * (P.gam - 1.0) / P.gam;
}

int initial_conditions(const Params &P, const grhd_config &c, std::vector<double> &out,
                       std::string &err)
{
    HostGrid G(P);
    out.assign((size_t)G.NXg * G.NYg * G.NZg * NEQ, 0.0);  // p = 0 (conditions.f90:34)
    const int nx = P.nx, ny = P.ny, nz = P.nz, ng = P.ng;
    auto X = [&](int i) { return xc(P, i); };
    auto Y = [&](int j) { return yc(P, j); };
    auto Z = [&](int k) { return zc(P, k); };
    const double pi = M_PI;
    switch (P.problem) {
    case 1: case 2: case 3: {  // shock tubes (conditions.f90:39-71)
        for (int i = -ng; i <= nx + ng; i++) {
            double rho, pr;
            if (P.problem == 1) {
                const bool L = X(i) < c.sod_interface;
                rho = L ? c.sod_density_left : c.sod_density_right;
                pr = L ? c.sod_pressure_left : c.sod_pressure_right;
            } else if (P.problem == 2) {
                const bool L = X(i) < c.strong_interface;
                rho = L ? c.strong_density_left : c.strong_density_right;
                pr = L ? c.strong_pressure_left : c.strong_pressure_right;
            } else {
                const bool L = X(i) < c.shu_interface;
                rho = L ? c.shu_density_left
                        : c.shu_density_right + c.shu_density_amplitude * std::sin(c.shu_wave_number * X(i));
                pr = L ? c.shu_pressure_left : c.shu_pressure_right;
            }
            for (int k = -ng; k <= nz + ng; k++)
                for (int j = -ng; j <= ny + ng; j++) {
                    out[G.at(EQ_DE, i, j, k)] = rho;
                    out[G.at(EQ_PR, i, j, k)] = pr;
                }
        }
        break;
    }
    case 4:  // Kelvin-Helmholtz (conditions.f90:73-103)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    const double xp = X(i), zp = Z(k);
                    const bool in = std::fabs(zp) < c.khi_half_width;
                    out[G.at(EQ_DE, i, j, k)] = in ? c.khi_density_inner : c.khi_density_outer;
                    out[G.at(EQ_VX, i, j, k)] = in ? c.khi_velocity_inner : c.khi_velocity_outer;
                    out[G.at(EQ_PR, i, j, k)] = c.khi_pressure;
                    out[G.at(EQ_VY, i, j, k)] = 0.0;
                    const double d = std::fabs(std::fabs(zp) - c.khi_half_width);
                    const double env = std::exp(-0.5 * ((d / c.khi_perturbation_width) * (d / c.khi_perturbation_width)));
                    out[G.at(EQ_VZ, i, j, k)] =
                        c.khi_perturbation_amplitude * std::sin(c.khi_perturbation_wave_number * pi * xp) * env;
                }
        break;
    case 5:  // jet (conditions.f90:105-124)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    const bool noz = X(i) <= c.jet_nozzle_radius && Z(k) <= c.jet_nozzle_length;
                    out[G.at(EQ_DE, i, j, k)] = noz ? c.jet_density : c.jet_ambient_density;
                    out[G.at(EQ_PR, i, j, k)] = noz ? c.jet_pressure : c.jet_ambient_pressure;
                    out[G.at(EQ_VZ, i, j, k)] = noz ? c.jet_velocity : c.jet_ambient_velocity;
                }
        break;
    case 6: {  // smooth advected wave (conditions.f90:126-143)
        const double arg = 0.5 * c.convergence_wave_number * pi * P.dx;
        const double sinc = (std::fabs(arg) > DBL_EPSILON) ? std::sin(arg) / arg : 1.0;
        for (int k = -ng; k <= nz + ng; k++)
            for (int j = -ng; j <= ny + ng; j++)
                for (int i = -ng; i <= nx + ng; i++) {
                    if (i >= 1 && i <= nx)
                        out[G.at(EQ_DE, i, j, k)] =
                            c.convergence_density +
                            c.convergence_amplitude * sinc * std::sin(c.convergence_wave_number * pi * X(i));
                    out[G.at(EQ_VX, i, j, k)] = c.convergence_velocity;
                    out[G.at(EQ_PR, i, j, k)] = c.convergence_pressure;
                }
        break;
    }
    case 7: {  // Michel accretion (conditions.f90:284-410), cells 1..nx+ng
        const double M = P.mass, gam = P.gam, g1 = P.g1;
        const double rc = c.michel_critical_radius, rhoc = c.michel_critical_density;
        const double uc = -std::sqrt(M / (2.0 * rc));
        const double utc = std::sqrt(1.0 - 2.0 * M / rc + uc * uc);
        const double Vc2 = (uc / utc) * (uc / utc);
        double pres = (Vc2 * rhoc) / (gam - Vc2 * g1);
        const double Kp = pres / std::pow(rhoc, gam);
        const double hs = 1.0 + g1 * pres / rhoc;
        const double C1 = rc * rc * rhoc * uc;
        const double C2 = rc * rc * rhoc * hs * uc * utc;
        const double eps = 1.0e-6;
        for (int i = 1; i <= nx + ng; i++) {
            const double r = log_radius(P) ? std::exp(X(i)) : X(i);
            double ur = -std::sqrt(2.0 * M / r);
            for (int it = 1; it <= 40; it++) {
                const double ut = std::sqrt(1.0 - 2.0 * M / r + ur * ur);
                double rho = C1 / (r * r * ur);
                double p_ = Kp * std::pow(rho, gam);
                double h = 1.0 + g1 * p_ / rho;
                const double f = h * ut - C2 / C1;
                const double urp = ur + eps, urm = ur - eps;
                const double utp = std::sqrt(1.0 - 2.0 * M / r + urp * urp);
                const double rp = C1 / (r * r * urp), pp = Kp * std::pow(rp, gam);
                const double fp = (1.0 + g1 * pp / rp) * utp - C2 / C1;
                const double utm = std::sqrt(1.0 - 2.0 * M / r + urm * urm);
                const double rm = C1 / (r * r * urm), pm = Kp * std::pow(rm, gam);
                const double fm = (1.0 + g1 * pm / rm) * utm - C2 / C1;
                const double df = (fp - fm) / (2.0 * eps);
                ur = ur - f / df;
                if (std::fabs(f) < 1.0e-12) break;
            }
            const double rho = C1 / (r * r * ur);
            pres = Kp * std::pow(rho, gam);
            for (int k = 1; k <= nz; k++)
                for (int j = 1; j <= ny; j++) {
                    FaceMetric m;
                    metric_any(P, X(i), Y(j), m);
                    const double ux = log_radius(P) ? ur / r : ur;
                    double g_tt = -(m.alpha * m.alpha);
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) g_tt = g_tt + m.g[a][b] * m.beta[a] * m.beta[b];
                    double g_tx = 0.0;
                    for (int b = 0; b < 3; b++) g_tx = g_tx + m.g[0][b] * m.beta[b];
                    const double A = g_tt, B = 2.0 * g_tx * ux, C = m.g[0][0] * (ux * ux) + 1.0;
                    const double ut = (-B - std::sqrt(B * B - 4.0 * A * C)) / (2.0 * A);
                    const double W = m.alpha * ut;
                    out[G.at(EQ_DE, i, j, k)] = rho;
                    out[G.at(EQ_PR, i, j, k)] = pres;
                    out[G.at(EQ_VX, i, j, k)] = ux / W + m.beta[0] / m.alpha;
                    out[G.at(EQ_VY, i, j, k)] = m.beta[1] / m.alpha;
                    out[G.at(EQ_VZ, i, j, k)] = m.beta[2] / m.alpha;
                }
        }
        break;
    }
    case 8: {  // dust free fall (conditions.f90:151-162, 249-278)
        for (int i = 1; i <= nx + ng; i++) {
            double r, jac;
            if (log_radius(P)) { r = std::exp(X(i)); jac = r; } else { r = X(i); jac = 1.0; }
            const double M = P.mass, ff = std::sqrt(2.0 * M / r);
            double q[NEQ] = {0, 0, 0, 0, 0};
            q[EQ_DE] = -c.dust_accretion_constant / (r * r * ff);
            q[EQ_PR] = P.p_floor;
            q[EQ_VX] = -1.0 / (std::sqrt(1.0 + r / (2.0 * M)) * (1.0 + ff + 2.0 * M / r)) / jac;
            for (int k = -ng; k <= nz + ng; k++)
                for (int j = -ng; j <= ny + ng; j++)
                    for (int v = 0; v < NEQ; v++) out[G.at(v, i, j, k)] = q[v];
        }
        for (int i = -ng; i <= 0; i++)
            for (int k = -ng; k <= nz + ng; k++)
                for (int j = -ng; j <= ny + ng; j++)
                    for (int v = 0; v < NEQ; v++) out[G.at(v, i, j, k)] = out[G.at(v, 1, j, k)];
        break;
    }
    case 9:  // off-axis blast (conditions.f90:164-190)
        for (int k = -ng; k <= nz + ng; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = -ng; i <= nx + ng; i++) {
                    const double r = log_radius(P) ? std::exp(X(i)) : X(i);
                    const double d2 = r * r + c.offaxis_radial_center * c.offaxis_radial_center -
                                      2.0 * c.offaxis_radial_center * r * std::cos(Z(k) - c.offaxis_phi_center);
                    const double ex = -d2 / (c.offaxis_width * c.offaxis_width);
                    const double gp = (ex < -40.0) ? 0.0 : std::exp(ex);
                    out[G.at(EQ_DE, i, j, k)] = c.offaxis_background_density + c.offaxis_density_amplitude * gp;
                    out[G.at(EQ_PR, i, j, k)] = c.offaxis_background_pressure + c.offaxis_pressure_amplitude * gp;
                }
        break;
    case 10: case 11: case 12: {  // Fishbone-Moncrief torus (conditions.f90:415-465)
        const double r_in = c.fm_inner_radius * P.mass, r_c = c.fm_pressure_max_radius * P.mass;
        double l_ang, W_in;
        if (!fishbone_constants(P, r_in, r_c, l_ang, W_in)) {
            err = "invalid Fishbone-Moncrief model parameters";
            return GRHD_ERR_INVALID_CONFIG;
        }
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    const double r = log_radius(P) ? std::exp(X(i)) : X(i);
                    FaceMetric m;
                    metric_any(P, X(i), Y(j), m);
                    double q[NEQ];
                    fishbone_state(P, r, r_in, c.fm_polytropic_constant, l_ang, W_in, m, q);
                    for (int v = 0; v < NEQ; v++) out[G.at(v, i, j, k)] = q[v];
                }
        break;
    }
    default:
        err = "unknown problem id";
        return GRHD_ERR_INVALID_CONFIG;
    }
    return GRHD_OK;
}

// Counter-based uniform deviate in [0, 1): splitmix64 of (seed, cell index).
static double uniform01(uint64_t seed, uint64_t idx)
{
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + idx + 0x632BE59BD9B4E5ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

void perturb(const Params &P, int type, double amplitude, uint64_t seed, std::vector<double> &p)
{
    // inject_pressure_noise / inject_density_noise (conditions.f90:603-636, 730-768):
    // multiplicative white noise on the cells denser than 5 % of the maximum.
    // The deviate comes from this counter-based generator (the reference uses the
    // compiler runtime's random_number, whose stream is not portable).  Type 3
    // (no reference counterpart) adds white noise A (2 U - 1) to v^z on the same
    // cells: the random part of the Kelvin-Helmholtz seed perturbation, whose
    // sinusoidal part the initial data carries (conditions.f90:73-103).
    HostGrid G(P);
    double rho_max = -1.0;
    for (int k = 1; k <= P.nz; k++)
        for (int j = 1; j <= P.ny; j++)
            for (int i = 1; i <= P.nx; i++) rho_max = std::fmax(rho_max, p[G.at(EQ_DE, i, j, k)]);
    for (int k = 1; k <= P.nz; k++)
        for (int j = 1; j <= P.ny; j++)
            for (int i = 1; i <= P.nx; i++) {
                if (!(p[G.at(EQ_DE, i, j, k)] > 0.05 * rho_max)) continue;
                const uint64_t idx = (((uint64_t)(k - 1) * P.ny) + (j - 1)) * P.nx + (i - 1);
                const double f = 1.0 + amplitude * (2.0 * uniform01(seed, idx) - 1.0);
                if (type == 1) {
                    p[G.at(EQ_PR, i, j, k)] *= f;
                } else if (type == 3) {
                    p[G.at(EQ_VZ, i, j, k)] += f - 1.0;
                } else {
                    p[G.at(EQ_DE, i, j, k)] *= f;
                    p[G.at(EQ_PR, i, j, k)] *= std::pow(f, P.gam);
                }
            }
}

}  // namespace grhd
