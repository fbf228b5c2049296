// grhd_initial.hpp -- host-side initial data and perturbations.
#pragma once
#include <string>
#include <vector>

#include "../../include/grhd_c.h"
#include "grhd_physics.hpp"

namespace grhd {

// Fills `out` with the full ghosted primitive array (reference layout) of the
// problem's initial state.  Returns a GRHD_* status.
int initial_conditions(const Params &P, const grhd_config &c, std::vector<double> &out,
                       std::string &err);
// l and W_in of the Fishbone-Moncrief torus (prepare_fishbone_moncrief).
bool fishbone_constants(const Params &P, double r_in, double r_max_dens, double &l_ang, double &W_in);
// Polytropic constant that puts rho = 1 at the pressure maximum.
double fishbone_K_for_unit_density(const Params &P, double r_in, double r_c);
// Multiplicative noise: type 1 pressure, 2 density (isentropic).
void perturb(const Params &P, int type, double amplitude, uint64_t seed, std::vector<double> &p);

}  // namespace grhd
