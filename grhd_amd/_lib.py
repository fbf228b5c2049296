"""ctypes binding of libgrhd_hip.so (include/grhd_c.h).

The library is loaded from the package directory.  There is no fallback: if it
is missing or fails to load, :func:`lib` raises, so no caller can silently run
on anything but the HIP kernels.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libgrhd_hip.so"

GRHD_OK = 0
GRHD_ERR_INVALID_METRIC = 2
ERRORS = {
    1: "invalid configuration",
    2: "invalid metric / CFL",
    3: "device out of memory",
    4: "HIP error",
    5: "invalid argument",
    6: "unsupported",
}


class GrhdError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERRORS.get(code, 'error')} ({code}): {message}")
        self.code = code


_D = ctypes.c_double
_I = ctypes.c_int32


class Config(ctypes.Structure):
    """Mirror of ``grhd_config`` (field order and types must match the header)."""

    _fields_ = [
        ("problem", _I), ("metric", _I), ("geometry", _I), ("log_r", _I),
        ("adiabatic_index", _D), ("bh_mass", _D), ("spin", _D),
        ("nx", _I), ("ny", _I), ("nz", _I), ("reconstruction", _I),
        ("r_min", _D), ("r_max", _D), ("y_min", _D), ("y_max", _D), ("z_min", _D), ("z_max", _D),
        ("riemann_solver", _I), ("tvd_limiter", _I), ("shock_sensor", _I), ("precision", _I),
        ("cfl", _D), ("final_time", _D),
        ("sod_interface", _D), ("sod_density_left", _D), ("sod_density_right", _D),
        ("sod_pressure_left", _D), ("sod_pressure_right", _D),
        ("strong_interface", _D), ("strong_density_left", _D), ("strong_density_right", _D),
        ("strong_pressure_left", _D), ("strong_pressure_right", _D),
        ("shu_interface", _D), ("shu_density_left", _D), ("shu_pressure_left", _D),
        ("shu_density_right", _D), ("shu_density_amplitude", _D), ("shu_wave_number", _D),
        ("shu_pressure_right", _D),
        ("khi_half_width", _D), ("khi_density_inner", _D), ("khi_density_outer", _D),
        ("khi_velocity_inner", _D), ("khi_velocity_outer", _D), ("khi_pressure", _D),
        ("khi_perturbation_amplitude", _D), ("khi_perturbation_wave_number", _D),
        ("khi_perturbation_width", _D),
        ("jet_ambient_density", _D), ("jet_ambient_pressure", _D), ("jet_ambient_velocity", _D),
        ("jet_nozzle_radius", _D), ("jet_nozzle_length", _D), ("jet_density", _D),
        ("jet_pressure", _D), ("jet_velocity", _D),
        ("convergence_density", _D), ("convergence_amplitude", _D), ("convergence_velocity", _D),
        ("convergence_pressure", _D), ("convergence_wave_number", _D),
        ("michel_critical_radius", _D), ("michel_critical_density", _D),
        ("dust_accretion_constant", _D),
        ("offaxis_radial_center", _D), ("offaxis_phi_center", _D), ("offaxis_width", _D),
        ("offaxis_background_density", _D), ("offaxis_background_pressure", _D),
        ("offaxis_density_amplitude", _D), ("offaxis_pressure_amplitude", _D),
        ("fm_inner_radius", _D), ("fm_pressure_max_radius", _D), ("fm_polytropic_constant", _D),
        ("allow_open_theta", _I), ("reserved", _I * 7),
    ]

    @classmethod
    def from_run(cls, cfg) -> "Config":
        """Build from a :class:`grhd_amd.params.RunConfig`."""
        c = cls()
        for name, _ in cls._fields_:
            if name == "reserved":
                continue
            if name == "allow_open_theta":
                c.allow_open_theta = 1 if cfg.extra.get("_allow_open_theta") else 0
                continue
            val = getattr(cfg, name)
            setattr(c, name, int(val) if isinstance(val, bool) else val)
        return c


_lib = None

_SIGS = {
    "grhd_create": ([ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "grhd_validate": ([ctypes.POINTER(Config)], ctypes.c_int),
    "grhd_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "grhd_last_error": ([ctypes.c_void_p], ctypes.c_char_p),
    "grhd_dims": ([ctypes.c_void_p, ctypes.POINTER(_I)], ctypes.c_int),
    "grhd_initial_conditions": ([ctypes.c_void_p], ctypes.c_int),
    "grhd_set_primitives": ([ctypes.c_void_p, ctypes.POINTER(_D)], ctypes.c_int),
    "grhd_perturb": ([ctypes.c_void_p, _I, _D, ctypes.c_uint64], ctypes.c_int),
    "grhd_fishbone_unit_density_k": ([ctypes.POINTER(Config), ctypes.POINTER(_D)], ctypes.c_int),
    "grhd_lapse_shift": ([ctypes.POINTER(Config), ctypes.POINTER(_D), ctypes.POINTER(_D)], ctypes.c_int),
    "grhd_put_primitives": ([ctypes.c_void_p, ctypes.POINTER(_D)], ctypes.c_int),
    "grhd_set_conserved": ([ctypes.c_void_p, ctypes.POINTER(_D)], ctypes.c_int),
    "grhd_put_conserved": ([ctypes.c_void_p, ctypes.POINTER(_D)], ctypes.c_int),
    "grhd_get_primitives": ([ctypes.c_void_p, ctypes.POINTER(_D)], ctypes.c_int),
    "grhd_get_conserved": ([ctypes.c_void_p, ctypes.POINTER(_D)], ctypes.c_int),
    "grhd_compute_dt": ([ctypes.c_void_p, ctypes.POINTER(_D)], ctypes.c_int),
    "grhd_step": ([ctypes.c_void_p, _D, _I], ctypes.c_int),
    "grhd_advance": ([ctypes.c_void_p, _D, _I, ctypes.POINTER(_D), ctypes.POINTER(_I),
                      ctypes.POINTER(_D)], ctypes.c_int),
    "grhd_boundaries": ([ctypes.c_void_p], ctypes.c_int),
    "grhd_rhs": ([ctypes.c_void_p, ctypes.POINTER(_D)], ctypes.c_int),
    "grhd_con2prim": ([ctypes.c_void_p], ctypes.c_int),
    "grhd_counters": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "grhd_total_mass": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "grhd_global_diagnostics": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "grhd_ppi_modes": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "grhd_accretion_rate": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    "grhd_convergence_norms": ([ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    "grhd_set_stream": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "grhd_synchronize": ([ctypes.c_void_p], ctypes.c_int),
    "grhd_last_timing": ([ctypes.c_void_p, ctypes.POINTER(_D), ctypes.POINTER(ctypes.c_int64)],
                         ctypes.c_int),
    "grhd_set_stage_form": ([ctypes.c_void_p, _I], ctypes.c_int),
    "grhd_set_face_form": ([ctypes.c_void_p, _I], ctypes.c_int),
    "grhd_footprint": ([ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
    "grhd_step_noalloc": ([ctypes.POINTER(Config), ctypes.c_void_p, ctypes.c_void_p, _D, _I, ctypes.c_void_p],
                          ctypes.c_int),
    "grhd_batch_create": ([_I, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "grhd_batch_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "grhd_batch_last_error": ([ctypes.c_void_p], ctypes.c_char_p),
    "grhd_batch_add": ([ctypes.c_void_p, ctypes.POINTER(Config), ctypes.POINTER(_I)], ctypes.c_int),
    "grhd_batch_problem": ([ctypes.c_void_p, _I, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "grhd_batch_remove": ([ctypes.c_void_p, _I], ctypes.c_int),
    "grhd_batch_step": ([ctypes.c_void_p, _I], ctypes.c_int),
    "grhd_batch_status": ([ctypes.c_void_p, _I, ctypes.POINTER(_I), ctypes.POINTER(_D),
                           ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "grhd_batch_set_variant": ([ctypes.c_void_p, _I], ctypes.c_int),
    "grhd_batch_set_pacing": ([ctypes.c_void_p, _I], ctypes.c_int),
    "grhd_batch_info": ([ctypes.c_void_p, ctypes.POINTER(_I), ctypes.POINTER(_I),
                         ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(_D),
                         ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
}


def declared_symbols(header: Path | None = None) -> list[str]:
    """Names of the functions declared in include/*.h."""
    import re
    root = Path(__file__).resolve().parent.parent / "include"
    names = []
    for h in ([header] if header else sorted(root.glob("*.h"))):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(grhd_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def lib():
    """Load libgrhd_hip.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(
                f"{LIB_PATH} is not built: run `python -m grhd_amd.build` (HIP path required)")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(rc: int, ctx=None) -> None:
    if rc != GRHD_OK:
        msg = lib().grhd_last_error(ctx)
        raise GrhdError(rc, msg.decode() if msg else "")
