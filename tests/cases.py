"""Parity cases: small .par files covering every metric kind, reconstruction,
Riemann solver, shock-sensor setting and boundary family of the update path.

Each is accepted by the reference's own parser and validation, so the same text
drives the reference (oracle/_ref), the C oracle and the HIP library.
"""

CASES = {
    # C1 exactly: relativistic Sod, WENO5 + HLLE, outflow (shock-tube BCs)
    "sod_c1": """problem = sod
adiabatic_index = 1.6666666666666667
nx = 400
sod_density_left = 10
sod_pressure_left = 13.333333333333334
sod_density_right = 1
sod_pressure_right = 6.666666666666667e-7
reconstruction = weno5
riemann_solver = hlle
cfl = 0.4
final_time = 0.4
""",
    "sod_mp5_hllc_sensor": """problem = sod
nx = 100
reconstruction = mp5
riemann_solver = hllc
shock_sensor = true
""",
    "strong_tvd_superbee_hllc": """problem = strong_shock
nx = 100
reconstruction = tvd
tvd_limiter = superbee
riemann_solver = hllc
""",
    "strong_godunov_hlle": """problem = strong_shock
nx = 80
reconstruction = godunov
""",
    "shu_weno3_sensor": """problem = shu_osher
nx = 160
reconstruction = weno3
shock_sensor = true
riemann_solver = hllc
""",
    "shu_tvd_minmod_sensor": """problem = shu_osher
nx = 120
reconstruction = tvd
tvd_limiter = minmod
shock_sensor = true
""",
    "conv_weno5_periodic": """problem = convergence
nx = 64
reconstruction = weno5
""",
    "khi_weno3_sensor": """problem = kelvin_helmholtz
nx = 32
nz = 32
reconstruction = weno3
shock_sensor = true
""",
    "jet_cylindrical": """problem = axisymmetric_jet
nx = 24
nz = 48
reconstruction = mp5
riemann_solver = hllc
""",
    "michel_ef_log": """problem = michel
logarithmic_r = true
nx = 128
nz = 1
r_min = 1.8
r_max = 20
michel_critical_radius = 8
""",
    "michel_ef_phys": """problem = michel
nx = 64
nz = 1
r_min = 1.8
r_max = 20
michel_critical_radius = 8
riemann_solver = hllc
""",
    # Michel accretion on Kerr-Schild spheroidal with spin, as the served
    # request mix draws it (initial data conditions.f90:284-410 on the KS
    # metric, outer injector conditions.f90:872-891)
    "michel_ks_a090": """problem = michel
metric = kerr_schild
geometry = spheroidal
logarithmic_r = true
spin = 0.9
nx = 96
nz = 1
r_min = 1.8
r_max = 20
michel_critical_radius = 8
""",
    "michel_ks_a099_hllc": """problem = michel
metric = kerr_schild
geometry = spheroidal
logarithmic_r = true
spin = 0.99
nx = 80
nz = 1
r_min = 1.8
r_max = 20
michel_critical_radius = 8
riemann_solver = hllc
""",
    "dust_ef_sensor": """problem = dust
nx = 64
nz = 4
""",
    "offaxis_ef": """problem = offaxis_blast
nx = 48
nz = 32
reconstruction = weno5
""",
    "offaxis_minkowski_spherical_log": """problem = offaxis_blast
metric = minkowski
geometry = spherical
bh_mass = 0
logarithmic_r = true
nx = 40
nz = 24
r_min = 2
r_max = 40
""",
    "offaxis_minkowski_spherical": """problem = offaxis_blast
metric = minkowski
geometry = spherical
bh_mass = 0
nx = 40
nz = 24
r_min = 2
r_max = 40
riemann_solver = hllc
""",
    "fm_equatorial_ks_log": """problem = fishbone_equatorial
nx = 48
nz = 16
fm_inner_radius = 6
fm_pressure_max_radius = 12
reconstruction = weno5
""",
    "fm_equatorial_ks_phys": """problem = fishbone_equatorial
logarithmic_r = false
nx = 48
nz = 16
r_min = 1.5
r_max = 30
fm_inner_radius = 6
fm_pressure_max_radius = 12
riemann_solver = hllc
""",
    "fm_sagittal_ef_log": """problem = fishbone_sagittal
nx = 32
ny = 24
r_min = 2
r_max = 40
reconstruction = mp5
""",
    # the headline kind (C4) at a reduced grid with theta = [0, pi]
    "fm3d_ks_mp5_hllc_sensor": """problem = fishbone_3d
metric = kerr_schild
geometry = spheroidal
logarithmic_r = true
spin = 0.9
nx = 24
ny = 16
nz = 8
r_min = 1.2
r_max = 40
fm_inner_radius = 6
fm_pressure_max_radius = 12
reconstruction = mp5
riemann_solver = hllc
shock_sensor = true
""",
    "fm3d_ks_weno5_hlle": """problem = fishbone_3d
metric = kerr_schild
geometry = spheroidal
logarithmic_r = true
spin = 0.9
nx = 20
ny = 12
nz = 8
r_min = 1.2
r_max = 40
fm_inner_radius = 6
fm_pressure_max_radius = 12
reconstruction = weno5
shock_sensor = false
""",
    "fm3d_ef": """problem = fishbone_3d
nx = 16
ny = 12
nz = 8
reconstruction = weno5
riemann_solver = hllc
""",
}

# cases whose initial state is perturbed (seeded) before stepping, so that
# smooth equilibria also exercise the non-trivial branches
NOISY = {"fm3d_ks_mp5_hllc_sensor": (2, 1.0e-2, 3435), "fm_equatorial_ks_log": (1, 1.0e-2, 3435),
         "michel_ef_log": (2, 1.0e-3, 3435), "michel_ks_a090": (2, 1.0e-3, 3435)}


# Michel accretion C2 (the reference's convergence case for the EF log-r
# metric): nr cells in r in [1.8, 20], critical radius 8, WENO5 + HLLE, to
# t = 500 with 1e-3 seeded density noise.
MICHEL_C2 = """problem = michel
logarithmic_r = true
nx = {nr}
nz = 1
r_min = 1.8
r_max = 20
michel_critical_radius = 8
reconstruction = weno5
riemann_solver = hlle
final_time = 500
"""


# Stationary dust (cold free fall) accretion onto Schwarzschild in EF
# coordinates, physical r in [1, 51]: the preset of problem 8 at nr cells on a
# single phi plane, to its preset final time t = 500.
DUST = """problem = dust
nx = {nr}
nz = 1
"""
