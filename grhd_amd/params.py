"""The GRHD2 ``.par`` operator surface: problem presets, parser, validation.

Mirrors the reference's configuration layer so that a ``.par`` file written for
the reference selects the same problem, metric, grid and numerical method here:

* presets: ``initialization.f90:483-821`` (one per problem id; the ``.par`` file
  has the last word, ``initialization.f90:444-446``);
* module defaults for keys no preset sets: ``variables.f90:111-243``;
* parser: ``parameters.f90:95-356`` (``name = value``, ``#`` comments,
  case-insensitive names and symbols, duplicate and unknown names are errors,
  theta/phi bounds in units of pi and possibly fractions ``a/b``,
  initial-condition names namespaced per problem);
* validation: ``parameters.f90:359-588``.

Errors raise :class:`ParameterError` (the reference stops with ``error stop 1``).
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field

PI = math.acos(-1.0)

# Identifiers (variables.f90:224-237) ------------------------------------------------
REC_GODUNOV, REC_TVD, REC_WENO3, REC_MP5, REC_WENO5 = 1, 2, 3, 4, 5
RS_HLLE, RS_HLLC = 1, 2
LIM_MINMOD, LIM_SUPERBEE, LIM_MC = 1, 2, 3
MINKOWSKI, EDDINGTON_FINKELSTEIN, KERR_SCHILD = 0, 1, 2
CARTESIAN, CYLINDRICAL, SPHERICAL, SPHEROIDAL = 0, 1, 2, 3
PRECISION_FP64, PRECISION_MIXED = 0, 1

PROBLEM_NAMES = {
    1: "sod", 2: "strong_shock", 3: "shu_osher", 4: "kelvin_helmholtz",
    5: "axisymmetric_jet", 6: "convergence", 7: "michel", 8: "dust",
    9: "offaxis_blast", 10: "fishbone_equatorial", 11: "fishbone_sagittal",
    12: "fishbone_3d",
}
_PROBLEM_ALIASES = {  # parameters.f90:893-921
    "sod": 1, "strong_shock": 2, "blast_wave": 2, "shu_osher": 3,
    "kelvin_helmholtz": 4, "khi": 4, "axisymmetric_jet": 5, "jet": 5,
    "convergence": 6, "convergence_test": 6, "michel": 7, "michel_accretion": 7,
    "dust": 8, "dust_accretion": 8, "offaxis_blast": 9, "off_axis_blast": 9,
    "fishbone_equatorial": 10, "fishbone_moncrief_equatorial": 10,
    "fishbone_sagittal": 11, "fishbone_sagital": 11, "fishbone_moncrief_sagittal": 11,
    "fishbone_3d": 12, "fishbone_moncrief_3d": 12,
}
_METRICS = {"minkowski": MINKOWSKI, "eddington_finkelstein": EDDINGTON_FINKELSTEIN,
            "ef": EDDINGTON_FINKELSTEIN, "kerr_schild": KERR_SCHILD, "ks": KERR_SCHILD}
_GEOMETRIES = {"cartesian": CARTESIAN, "cylindrical": CYLINDRICAL,
               "spherical": SPHERICAL, "spheroidal": SPHEROIDAL}
_RECON = {"godunov": REC_GODUNOV, "tvd": REC_TVD, "weno3": REC_WENO3, "mp5": REC_MP5,
          "weno5": REC_WENO5}
_RIEMANN = {"hlle": RS_HLLE, "hllc": RS_HLLC}
_LIMITERS = {"minmod": LIM_MINMOD, "superbee": LIM_SUPERBEE, "mc": LIM_MC}
_PERTURBATIONS = {"none": 0, "pressure_noise": 1, "density_noise": 2, "density_mode": 3}
_TRUE = {"1", "true", ".true.", "yes", "si"}
_FALSE = {"0", "false", ".false.", "no"}


class ParameterError(ValueError):
    """A ``.par`` file or configuration the reference would reject."""


@dataclass
class RunConfig:
    """A fully resolved run: preset, then ``.par`` overrides."""

    problem: int = 1
    metric: int = MINKOWSKI
    geometry: int = CARTESIAN
    log_r: bool = False
    adiabatic_index: float = 1.4
    bh_mass: float = 0.0
    spin: float = 0.0
    nx: int = 1
    ny: int = 1
    nz: int = 1
    r_min: float = 0.0
    r_max: float = 1.0
    y_min: float = 0.0
    y_max: float = 1.0
    z_min: float = 0.0
    z_max: float = 1.0
    reconstruction: int = REC_WENO5
    riemann_solver: int = RS_HLLE
    tvd_limiter: int = LIM_MC
    shock_sensor: bool = True
    cfl: float = 0.4
    final_time: float = 0.4
    precision: int = PRECISION_FP64
    # initial-condition parameters (variables.f90:166-219)
    sod_interface: float = 0.5
    sod_density_left: float = 1.0
    sod_density_right: float = 0.125
    sod_pressure_left: float = 1.0
    sod_pressure_right: float = 0.1
    strong_interface: float = 0.5
    strong_density_left: float = 1.0
    strong_density_right: float = 1.0
    strong_pressure_left: float = 1000.0
    strong_pressure_right: float = 0.01
    shu_interface: float = 0.5
    shu_density_left: float = 5.0
    shu_pressure_left: float = 50.0
    shu_density_right: float = 2.0
    shu_density_amplitude: float = 0.3
    shu_wave_number: float = 50.0
    shu_pressure_right: float = 5.0
    khi_half_width: float = 0.25
    khi_density_inner: float = 2.0
    khi_density_outer: float = 1.0
    khi_velocity_inner: float = 0.5
    khi_velocity_outer: float = -0.5
    khi_pressure: float = 2.5
    khi_perturbation_amplitude: float = 0.01
    khi_perturbation_wave_number: float = 10.0
    khi_perturbation_width: float = 0.05
    jet_ambient_density: float = 10.0
    jet_ambient_pressure: float = 0.01
    jet_ambient_velocity: float = 0.0
    jet_nozzle_radius: float = 1.0
    jet_nozzle_length: float = 1.0
    jet_density: float = 0.1
    jet_pressure: float = 0.01
    jet_velocity: float = 0.99
    convergence_density: float = 1.0
    convergence_amplitude: float = 0.1
    convergence_velocity: float = 0.5
    convergence_pressure: float = 1.0
    convergence_wave_number: float = 2.0
    michel_critical_radius: float = 100.0
    michel_critical_density: float = 0.1
    dust_accretion_constant: float = -0.5
    offaxis_radial_center: float = 15.0
    offaxis_phi_center: float = PI
    offaxis_width: float = 1.5
    offaxis_background_density: float = 0.1
    offaxis_background_pressure: float = 0.1
    offaxis_density_amplitude: float = 1.0
    offaxis_pressure_amplitude: float = 1.0
    fm_inner_radius: float = 6.25
    fm_pressure_max_radius: float = 9.25
    fm_polytropic_constant: float = 0.0015
    # run control the update path does not read, kept for the operator surface
    save_interval: float = 1.0
    checkpoint_interval: float = 100.0
    output_prefix: str = "out"
    output_folder: str = "out"
    perturbation: int = 0
    perturbation_seed: int = 3435
    perturbation_time: float = 1000.0
    perturbation_amplitude: float = 0.01
    perturbation_mode: float = 1.0
    extra: dict = field(default_factory=dict)

    @property
    def nghost(self) -> int:
        """choose_numerical_architecture (initialization.f90:26-55)."""
        return {REC_GODUNOV: 1, REC_TVD: 2, REC_WENO3: 2}.get(self.reconstruction, 3)

    @property
    def n_cells(self) -> int:
        return self.nx * self.ny * self.nz

    def replace(self, **kw) -> "RunConfig":
        return dataclasses.replace(self, **kw)


def preset(problem: int) -> RunConfig:
    """Problem presets (initialization.f90:483-821) on top of module defaults."""
    c = RunConfig(problem=problem)
    two_pi = 2.0 * PI
    if problem == 1:  # setup_sod 483-508
        c = c.replace(geometry=CARTESIAN, metric=MINKOWSKI, shock_sensor=False,
                      adiabatic_index=1.4, nx=200, r_min=0.0, r_max=1.0,
                      final_time=0.4, cfl=0.4, save_interval=0.1)
    elif problem == 2:  # 510-535
        c = c.replace(geometry=CARTESIAN, metric=MINKOWSKI, shock_sensor=False,
                      adiabatic_index=5.0 / 3.0, nx=400, r_min=0.0, r_max=1.0,
                      final_time=0.4, cfl=0.4, save_interval=0.1)
    elif problem == 3:  # 537-564
        c = c.replace(geometry=CARTESIAN, metric=MINKOWSKI, shock_sensor=False,
                      adiabatic_index=5.0 / 3.0, nx=400, r_min=0.0, r_max=1.0,
                      final_time=0.4, cfl=0.4, save_interval=0.1)
    elif problem == 4:  # 566-595
        c = c.replace(geometry=CARTESIAN, metric=MINKOWSKI, shock_sensor=False,
                      adiabatic_index=5.0 / 3.0, nx=400, r_min=-0.5, r_max=0.5,
                      ny=1, y_min=0.0, y_max=1.0, nz=400, z_min=-0.5, z_max=0.5,
                      final_time=3.0, cfl=0.25, save_interval=0.1)
    elif problem == 5:  # 597-627
        c = c.replace(geometry=CYLINDRICAL, metric=MINKOWSKI, shock_sensor=True,
                      adiabatic_index=5.0 / 3.0, nx=320, r_min=0.0, r_max=16.0,
                      ny=1, y_min=0.0, y_max=PI, nz=800, z_min=0.0, z_max=40.0,
                      final_time=120.0, cfl=0.25, save_interval=1.0)
    elif problem == 6:  # 629-654
        c = c.replace(geometry=CARTESIAN, metric=MINKOWSKI, shock_sensor=False,
                      adiabatic_index=4.0 / 3.0, nx=2048, r_min=0.0, r_max=1.0,
                      final_time=0.7, cfl=0.10, save_interval=0.1)
    elif problem == 7:  # 656-682
        c = c.replace(geometry=SPHERICAL, metric=EDDINGTON_FINKELSTEIN, shock_sensor=False,
                      adiabatic_index=4.0 / 3.0, bh_mass=1.0, nx=400, r_min=1.01,
                      r_max=51.0, ny=1, y_min=0.0, y_max=PI, nz=200, z_min=0.0,
                      z_max=two_pi, final_time=500.0, cfl=0.5, save_interval=1.0)
    elif problem == 8:  # 684-706
        c = c.replace(geometry=SPHERICAL, metric=EDDINGTON_FINKELSTEIN, shock_sensor=True,
                      adiabatic_index=4.0 / 3.0, bh_mass=1.0, nx=500, r_min=1.0,
                      r_max=51.0, ny=1, y_min=0.0, y_max=PI, nz=4, z_min=0.0,
                      z_max=two_pi, final_time=500.0, cfl=0.5, save_interval=50.0)
    elif problem == 9:  # 708-736
        c = c.replace(geometry=SPHERICAL, metric=EDDINGTON_FINKELSTEIN, shock_sensor=True,
                      adiabatic_index=4.0 / 3.0, bh_mass=1.0, nx=400, r_min=1.0,
                      r_max=121.0, ny=1, y_min=0.0, y_max=PI, nz=200, z_min=0.0,
                      z_max=two_pi, final_time=2000.0, cfl=0.4, save_interval=25.0)
    elif problem == 10:  # 738-766
        c = c.replace(geometry=SPHEROIDAL, metric=KERR_SCHILD, shock_sensor=False,
                      log_r=True, adiabatic_index=4.0 / 3.0, bh_mass=1.0, spin=0.9,
                      fm_inner_radius=6.25, fm_pressure_max_radius=9.25,
                      fm_polytropic_constant=0.0015, nx=400, r_min=1.2, r_max=40.0,
                      ny=1, y_min=0.0, y_max=PI, nz=200, z_min=0.0, z_max=two_pi,
                      final_time=1000.0, cfl=0.4, save_interval=10.0)
    elif problem == 11:  # 768-793
        c = c.replace(geometry=SPHERICAL, metric=EDDINGTON_FINKELSTEIN, shock_sensor=True,
                      log_r=True, adiabatic_index=4.0 / 3.0, bh_mass=1.0,
                      fm_inner_radius=6.25, fm_pressure_max_radius=9.25,
                      fm_polytropic_constant=0.0015, nx=400, r_min=2.0, r_max=200.0,
                      ny=200, y_min=0.0, y_max=PI, nz=1, z_min=0.0, z_max=two_pi,
                      final_time=0.0, cfl=0.4, save_interval=1.0)
    elif problem == 12:  # 795-821
        c = c.replace(geometry=SPHERICAL, metric=EDDINGTON_FINKELSTEIN, shock_sensor=True,
                      adiabatic_index=4.0 / 3.0, bh_mass=1.0, fm_inner_radius=6.0,
                      fm_pressure_max_radius=12.0, fm_polytropic_constant=0.015,
                      nx=100, r_min=1.0, r_max=121.0, ny=100, y_min=0.0, y_max=PI,
                      nz=100, z_min=0.0, z_max=two_pi, final_time=35.0, cfl=0.4,
                      save_interval=1.0)
    else:
        raise ParameterError(f"invalid problem id {problem}")
    return c


# parameters.f90 field map: .par key -> (RunConfig attribute, parser)
def _real(text: str) -> float:
    try:
        return float(text.replace("d", "e").replace("D", "e"))
    except ValueError as exc:
        raise ParameterError(f"expected a real number: {text}") from exc


def _int(text: str) -> int:
    try:
        return int(text)
    except ValueError as exc:
        raise ParameterError(f"expected an integer: {text}") from exc


def _logical(text: str) -> bool:
    t = text.strip().lower()
    if t in _TRUE:
        return True
    if t in _FALSE:
        return False
    raise ParameterError(f"expected true/false, yes/no or 1/0: {text}")


def _pi_coeff(text: str) -> float:
    """parse_pi_coefficient (parameters.f90:808-835)."""
    t = text.strip()
    if "/" not in t:
        return _real(t)
    num, _, den = t.partition("/")
    if not num or not den or "/" in den:
        raise ParameterError(f"invalid fraction of pi: {text}")
    d = _real(den)
    if d == 0.0:
        raise ParameterError(f"invalid or zero denominator: {text}")
    return _real(num) / d


def _string(text: str) -> str:
    t = text.strip()
    if len(t) >= 2 and t[0] == t[-1] and t[0] in "\"'":
        t = t[1:-1]
    if not t:
        raise ParameterError("string values cannot be empty")
    return t


def _symbol(table: dict, what: str):
    def parse(text: str):
        t = text.strip().lower()
        if t not in table:
            raise ParameterError(f"unknown {what}: {text}")
        return table[t]
    return parse


_REAL_KEYS = {
    "bh_mass": "bh_mass", "spin": "spin", "adiabatic_index": "adiabatic_index",
    "r_min": "r_min", "r_max": "r_max", "y_min": "y_min", "y_max": "y_max",
    "z_min": "z_min", "z_max": "z_max", "final_time": "final_time", "cfl": "cfl",
    "save_interval": "save_interval", "checkpoint_interval": "checkpoint_interval",
    "perturbation_time": "perturbation_time",
    "perturbation_amplitude": "perturbation_amplitude",
    "perturbation_mode": "perturbation_mode",
}
_IC_REAL_KEYS = [
    "sod_interface", "sod_density_left", "sod_density_right", "sod_pressure_left",
    "sod_pressure_right", "strong_interface", "strong_density_left",
    "strong_density_right", "strong_pressure_left", "strong_pressure_right",
    "shu_interface", "shu_density_left", "shu_pressure_left", "shu_density_right",
    "shu_density_amplitude", "shu_wave_number", "shu_pressure_right", "khi_half_width",
    "khi_density_inner", "khi_density_outer", "khi_velocity_inner", "khi_velocity_outer",
    "khi_pressure", "khi_perturbation_amplitude", "khi_perturbation_wave_number",
    "khi_perturbation_width", "jet_ambient_density", "jet_ambient_pressure",
    "jet_ambient_velocity", "jet_nozzle_radius", "jet_nozzle_length", "jet_density",
    "jet_pressure", "jet_velocity", "convergence_density", "convergence_amplitude",
    "convergence_velocity", "convergence_pressure", "michel_critical_radius",
    "michel_critical_density", "dust_accretion_constant", "offaxis_radial_center",
    "offaxis_width", "offaxis_background_density", "offaxis_background_pressure",
    "offaxis_density_amplitude", "offaxis_pressure_amplitude", "fm_inner_radius",
    "fm_pressure_max_radius", "fm_polytropic_constant",
]
_IC_PREFIX = {1: "sod_", 2: "strong_", 3: "shu_", 4: "khi_", 5: "jet_", 6: "convergence_",
              7: "michel_", 8: "dust_", 9: "offaxis_", 10: "fm_", 11: "fm_", 12: "fm_"}
_RESTART_CONTROL = {  # parameters.f90:716-731
    "restart", "restart_file", "final_time", "cfl", "save_interval", "checkpoint_interval",
    "output_prefix", "output_folder", "vtk_mapping", "shock_sensor", "extract_gw",
    "extract_mdot", "ppi_diagnostics", "diagnostic_stride", "extraction_stride",
    "mass_monitor_stride", "perturbation", "perturbation_seed", "perturbation_time",
    "perturbation_amplitude", "perturbation_mode",
}
_PASSTHROUGH = {"restart_file", "vtk_mapping", "extract_gw", "extract_mdot", "ppi_diagnostics",
                "diagnostic_stride", "extraction_stride", "mass_monitor_stride", "restart"}


def parse_par_lines(text: str, source: str = "<par>") -> list[tuple[int, str, str]]:
    """Tokenise a .par text into (line, key, value) (parameters.f90:114-135)."""
    out, seen = [], set()
    for lineno, raw in enumerate(text.splitlines(), start=1):
        line = raw.split("#", 1)[0]
        if not line.strip():
            continue
        eq = line.find("=")
        if eq <= 0 or eq == len(line.rstrip()) - 1:
            raise ParameterError(f"{source}:{lineno}: expected name = value")
        if "=" in line[eq + 1:]:
            raise ParameterError(f"{source}:{lineno}: more than one '='")
        key = line[:eq].strip().lower()
        value = line[eq + 1:].strip()
        if not key or len(key) > 64:
            raise ParameterError(f"{source}:{lineno}: empty or too long parameter name")
        if key in seen:
            raise ParameterError(f"{source}:{lineno}: duplicated parameter: {key}")
        seen.add(key)
        out.append((lineno, key, value))
    return out


def config_from_par(text: str, source: str = "<par>") -> RunConfig:
    """Cold-start configuration: preset of ``problem``, then the file's keys."""
    entries = parse_par_lines(text, source)
    problem = None
    for _, key, value in entries:
        if key == "problem":
            problem = _symbol(_PROBLEM_ALIASES, "problem")(value)
    if problem is None:
        raise ParameterError(f"{source}: missing mandatory parameter 'problem' for a cold start")
    cfg = preset(problem)
    for lineno, key, value in entries:
        where = f"{source}:{lineno}"
        try:
            _apply_key(cfg, key, value, problem)
        except ParameterError as exc:
            raise ParameterError(f"{where}: {exc}") from None
    validate(cfg)
    return cfg


def _apply_key(cfg: RunConfig, key: str, value: str, problem: int) -> None:
    if key in ("problem",):
        return
    if any(key.startswith(p) for p in set(_IC_PREFIX.values())) and \
            not key.startswith(_IC_PREFIX[problem]):
        raise ParameterError(f"parameter does not belong to the selected problem: {key}")
    if key in _REAL_KEYS:
        setattr(cfg, _REAL_KEYS[key], _real(value))
    elif key in _IC_REAL_KEYS:
        setattr(cfg, key, _real(value))
    elif key == "convergence_wave_number":
        cfg.convergence_wave_number = _pi_coeff(value)
    elif key == "offaxis_phi_center":
        cfg.offaxis_phi_center = _pi_coeff(value) * PI
    elif key in ("nx", "ny", "nz", "perturbation_seed"):
        setattr(cfg, key, _int(value))
    elif key == "theta_min":
        cfg.y_min = _pi_coeff(value) * PI
    elif key == "theta_max":
        cfg.y_max = _pi_coeff(value) * PI
    elif key == "phi_min":
        cfg.z_min = _pi_coeff(value) * PI
    elif key == "phi_max":
        cfg.z_max = _pi_coeff(value) * PI
    elif key == "metric":
        cfg.metric = _symbol(_METRICS, "metric")(value)
    elif key == "geometry":
        cfg.geometry = _symbol(_GEOMETRIES, "geometry")(value)
    elif key == "logarithmic_r":
        cfg.log_r = _logical(value)
    elif key == "reconstruction":
        cfg.reconstruction = _symbol(_RECON, "reconstruction")(value)
    elif key == "riemann_solver":
        cfg.riemann_solver = _symbol(_RIEMANN, "riemann_solver")(value)
    elif key == "tvd_limiter":
        cfg.tvd_limiter = _symbol(_LIMITERS, "tvd_limiter")(value)
    elif key == "shock_sensor":
        cfg.shock_sensor = _logical(value)
    elif key == "precision":  # build extension: fp64 | mixed
        cfg.precision = _symbol({"fp64": PRECISION_FP64, "mixed": PRECISION_MIXED},
                                "precision")(value)
    elif key in ("output_prefix", "output_folder"):
        setattr(cfg, key, _string(value))
    elif key == "perturbation":
        cfg.perturbation = _symbol(_PERTURBATIONS, "perturbation")(value)
    elif key in _PASSTHROUGH:
        cfg.extra[key] = value
    else:
        raise ParameterError(f"unknown parameter: {key}")


def validate(cfg: RunConfig) -> None:
    """validate_parameters + validate_initial_parameters (parameters.f90:359-588)."""
    eps = 2.220446049250313e-16
    reals = [cfg.r_min, cfg.r_max, cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max,
             cfg.adiabatic_index, cfg.bh_mass, cfg.spin, cfg.final_time, cfg.cfl]
    if not all(math.isfinite(v) for v in reals):
        raise ParameterError("real parameters must be finite")
    polar = cfg.geometry in (SPHERICAL, SPHEROIDAL)
    if cfg.nx < 1 or cfg.ny < 1 or cfg.nz < 1:
        raise ParameterError("nx, ny and nz must be positive")
    if cfg.r_max <= cfg.r_min:
        raise ParameterError("r_max must exceed r_min")
    if cfg.log_r and not polar:
        raise ParameterError("a logarithmic radial grid needs spherical or spheroidal geometry")
    if polar and cfg.r_min <= 0.0:
        raise ParameterError("spherical and spheroidal geometries need r_min > 0")
    if cfg.adiabatic_index <= 1.0:
        raise ParameterError("adiabatic_index must exceed 1")
    if cfg.bh_mass < 0.0:
        raise ParameterError("bh_mass cannot be negative")
    if cfg.final_time < 0.0 or cfg.cfl <= 0.0:
        raise ParameterError("need final_time >= 0 and cfl > 0")
    if cfg.metric == KERR_SCHILD and cfg.geometry == SPHERICAL:
        cfg.geometry = SPHEROIDAL  # parameters.f90:428-431
    if cfg.metric == MINKOWSKI:
        if abs(cfg.spin) > 0.0 or abs(cfg.bh_mass) > 0.0:
            raise ParameterError("Minkowski needs spin = 0 and bh_mass = 0")
        if cfg.geometry not in (CARTESIAN, CYLINDRICAL, SPHERICAL):
            raise ParameterError("Minkowski needs cartesian, cylindrical or spherical geometry")
    elif cfg.metric == EDDINGTON_FINKELSTEIN:
        if abs(cfg.spin) > 0.0:
            raise ParameterError("Eddington-Finkelstein needs spin = 0")
        if cfg.bh_mass <= 0.0 or cfg.geometry != SPHERICAL:
            raise ParameterError("Eddington-Finkelstein needs bh_mass > 0 and spherical geometry")
    else:
        if cfg.bh_mass <= 0.0 or cfg.geometry != SPHEROIDAL:
            raise ParameterError("Kerr-Schild needs bh_mass > 0 and spheroidal geometry")
        if abs(cfg.spin) > cfg.bh_mass + 64.0 * eps * max(1.0, cfg.bh_mass):
            raise ParameterError("Kerr-Schild needs |spin| <= bh_mass")
    if polar:
        if cfg.y_min < 0.0 or cfg.y_max > PI or cfg.y_max <= cfg.y_min:
            raise ParameterError("need 0 <= theta_min < theta_max <= 1 (units of pi)")
        if cfg.z_max <= cfg.z_min or cfg.z_max - cfg.z_min > 2.0 * PI + 64.0 * eps * PI:
            raise ParameterError("the azimuthal interval must have length in (0, 2 pi]")
    elif cfg.y_max <= cfg.y_min or cfg.z_max <= cfg.z_min:
        raise ParameterError("coordinate bounds must increase")
    _validate_initial(cfg)


def _validate_initial(cfg: RunConfig) -> None:
    """validate_initial_parameters (parameters.f90:488-588)."""
    tol = 128.0 * 2.220446049250313e-16 * PI
    p = cfg.problem
    bad = False
    if p == 1:
        bad = (cfg.sod_interface <= cfg.r_min or cfg.sod_interface >= cfg.r_max or
               min(cfg.sod_density_left, cfg.sod_density_right) <= 0.0 or
               min(cfg.sod_pressure_left, cfg.sod_pressure_right) <= 0.0)
    elif p == 2:
        bad = (cfg.strong_interface <= cfg.r_min or cfg.strong_interface >= cfg.r_max or
               min(cfg.strong_density_left, cfg.strong_density_right) <= 0.0 or
               min(cfg.strong_pressure_left, cfg.strong_pressure_right) <= 0.0)
    elif p == 3:
        bad = (cfg.shu_interface <= cfg.r_min or cfg.shu_interface >= cfg.r_max or
               cfg.shu_density_left <= 0.0 or
               cfg.shu_density_right <= abs(cfg.shu_density_amplitude) or
               min(cfg.shu_pressure_left, cfg.shu_pressure_right) <= 0.0 or
               cfg.shu_wave_number <= 0.0)
    elif p == 4:
        bad = (cfg.khi_half_width <= 0.0 or min(cfg.khi_density_inner, cfg.khi_density_outer) <= 0
               or cfg.khi_pressure <= 0.0 or abs(cfg.khi_velocity_inner) >= 1.0 or
               abs(cfg.khi_velocity_outer) >= 1.0 or cfg.khi_perturbation_amplitude < 0.0 or
               cfg.khi_perturbation_amplitude >= 1.0 or cfg.khi_perturbation_wave_number <= 0
               or cfg.khi_perturbation_width <= 0.0 or
               max(cfg.khi_velocity_inner ** 2, cfg.khi_velocity_outer ** 2) +
               cfg.khi_perturbation_amplitude ** 2 >= 1.0)
    elif p == 5:
        bad = (min(cfg.jet_ambient_density, cfg.jet_density) <= 0.0 or
               min(cfg.jet_ambient_pressure, cfg.jet_pressure) <= 0.0 or
               cfg.jet_nozzle_radius <= 0.0 or cfg.jet_nozzle_length <= 0.0 or
               abs(cfg.jet_ambient_velocity) >= 1.0 or abs(cfg.jet_velocity) >= 1.0)
    elif p == 6:
        bad = (cfg.convergence_density <= abs(cfg.convergence_amplitude) or
               cfg.convergence_pressure <= 0.0 or abs(cfg.convergence_velocity) >= 1.0 or
               cfg.convergence_wave_number <= 0.0)
    elif p == 7:
        bad = (cfg.bh_mass <= 0.0 or cfg.michel_critical_radius <= 1.5 * cfg.bh_mass or
               cfg.michel_critical_density <= 0.0)
    elif p == 8:
        bad = cfg.bh_mass <= 0.0 or cfg.dust_accretion_constant >= 0.0
    elif p == 9:
        bad = (cfg.offaxis_radial_center <= 0.0 or cfg.offaxis_width <= 0.0 or
               cfg.offaxis_background_density <= 0.0 or
               cfg.offaxis_background_density + cfg.offaxis_density_amplitude <= 0.0 or
               cfg.offaxis_background_pressure <= 0.0 or
               cfg.offaxis_background_pressure + cfg.offaxis_pressure_amplitude <= 0.0)
    elif p in (10, 11, 12):
        bad = (cfg.bh_mass <= 0.0 or cfg.fm_inner_radius <= 0.0 or
               cfg.fm_pressure_max_radius <= cfg.fm_inner_radius or
               cfg.fm_polytropic_constant <= 0.0)
        if not bad and p == 10:
            if cfg.ny != 1 or abs(0.5 * (cfg.y_min + cfg.y_max) - 0.5 * PI) > tol:
                raise ParameterError("fishbone_equatorial needs ny=1 centred on theta=pi/2")
        if not bad and p == 11:
            if cfg.ny <= 1 or cfg.nz != 1 or abs(cfg.y_min) > tol or abs(cfg.y_max - PI) > tol:
                raise ParameterError("fishbone_sagittal needs ny>1, nz=1 and theta=[0,1]")
        if not bad and p == 12:
            # with _allow_open_theta (a build option) theta may stop short of
            # the poles; such a boundary gets an outflow ghost fill
            open_ok = cfg.extra.get("_allow_open_theta")
            if (cfg.ny <= 1 or cfg.nz <= 1 or cfg.nz % 2 != 0 or
                    (not open_ok and (abs(cfg.y_min) > tol or abs(cfg.y_max - PI) > tol)) or
                    abs((cfg.z_max - cfg.z_min) - 2.0 * PI) > tol):
                raise ParameterError(
                    "fishbone_3d needs ny,nz>1, even nz, theta=[0,1] and phi length 2 [pi]")
    else:
        raise ParameterError("invalid problem id")
    if bad:
        raise ParameterError(f"invalid initial-condition parameters for {PROBLEM_NAMES[p]}")


def config_to_par(cfg: RunConfig, output_folder: str = "/tmp/grhd_out") -> str:
    """Write a .par text that makes the reference set up ``cfg`` (for the oracle)."""
    inv_metric = {MINKOWSKI: "minkowski", EDDINGTON_FINKELSTEIN: "eddington_finkelstein",
                  KERR_SCHILD: "kerr_schild"}
    inv_geom = {CARTESIAN: "cartesian", CYLINDRICAL: "cylindrical", SPHERICAL: "spherical",
                SPHEROIDAL: "spheroidal"}
    inv_rec = {v: k for k, v in _RECON.items()}
    inv_rs = {v: k for k, v in _RIEMANN.items()}
    inv_lim = {v: k for k, v in _LIMITERS.items()}
    polar = cfg.geometry in (SPHERICAL, SPHEROIDAL)
    lines = [
        f"problem = {PROBLEM_NAMES[cfg.problem]}",
        f"metric = {inv_metric[cfg.metric]}",
        f"geometry = {inv_geom[cfg.geometry]}",
        f"logarithmic_r = {'true' if cfg.log_r else 'false'}",
        f"adiabatic_index = {cfg.adiabatic_index!r}",
        f"bh_mass = {cfg.bh_mass!r}",
        f"spin = {cfg.spin!r}",
        f"nx = {cfg.nx}", f"ny = {cfg.ny}", f"nz = {cfg.nz}",
        f"r_min = {cfg.r_min!r}", f"r_max = {cfg.r_max!r}",
        f"reconstruction = {inv_rec[cfg.reconstruction]}",
        f"riemann_solver = {inv_rs[cfg.riemann_solver]}",
        f"tvd_limiter = {inv_lim[cfg.tvd_limiter]}",
        f"shock_sensor = {'true' if cfg.shock_sensor else 'false'}",
        f"cfl = {cfg.cfl!r}", f"final_time = {cfg.final_time!r}",
        f"output_folder = {output_folder}",
    ]
    if polar:
        lines += [f"theta_min = {cfg.y_min / PI!r}", f"theta_max = {cfg.y_max / PI!r}",
                  f"phi_min = {cfg.z_min / PI!r}", f"phi_max = {cfg.z_max / PI!r}"]
    else:
        lines += [f"y_min = {cfg.y_min!r}", f"y_max = {cfg.y_max!r}",
                  f"z_min = {cfg.z_min!r}", f"z_max = {cfg.z_max!r}"]
    prefix = _IC_PREFIX[cfg.problem]
    for key in _IC_REAL_KEYS:
        if key.startswith(prefix):
            lines.append(f"{key} = {getattr(cfg, key)!r}")
    if cfg.problem == 6:
        lines.append(f"convergence_wave_number = {cfg.convergence_wave_number!r}")
    if cfg.problem == 9:
        lines.append(f"offaxis_phi_center = {cfg.offaxis_phi_center / PI!r}")
    return "\n".join(lines) + "\n"
