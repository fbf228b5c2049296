"""Host-side mirror of the reference update path on top of the C ABI.

:class:`Solver` owns one device context.  Its methods are named after the
reference routines they replace; arrays use the reference layout (see
include/grhd_c.h): primitives ``(nz+2ng+1, ny+2ng+1, nx+2ng+1, 5)``, conserved
``(nz, ny, nx, 5)``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .params import RunConfig

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError("arrays must be C-contiguous float64")
    return a.ctypes.data_as(_dp)


def fishbone_unit_density_k(cfg: RunConfig) -> float:
    """Polytropic K of the torus with rho = 1 at the pressure maximum."""
    c = _lib.Config.from_run(cfg)
    k = ctypes.c_double()
    _lib.check(_lib.lib().grhd_fishbone_unit_density_k(ctypes.byref(c), ctypes.byref(k)))
    return k.value


def lapse_shift(cfg: RunConfig):
    """alpha (nz, ny, nx) and beta (nz, ny, nx, 3) at the interior cell
    centres, evaluated on the host (no device needed)."""
    c = _lib.Config.from_run(cfg)
    alpha = np.zeros((cfg.nz, cfg.ny, cfg.nx))
    beta = np.zeros((cfg.nz, cfg.ny, cfg.nx, 3))
    _lib.check(_lib.lib().grhd_lapse_shift(ctypes.byref(c), _ptr(alpha), _ptr(beta)))
    return alpha, beta


class Solver:
    def __init__(self, cfg: RunConfig, handle=None):
        """A new device context for cfg, or (handle given) a view of a context
        owned elsewhere -- a batch problem -- that close() leaves alive."""
        self.cfg = cfg
        self._L = _lib.lib()
        self._c = _lib.Config.from_run(cfg)
        self._owned = handle is None
        if handle is None:
            h = ctypes.c_void_p()
            _lib.check(self._L.grhd_create(ctypes.byref(self._c), ctypes.byref(h)))
        else:
            h = handle
        self.h = h
        dims = (ctypes.c_int32 * 4)()
        self._chk(self._L.grhd_dims(self.h, dims))
        self.nx, self.ny, self.nz, self.ng = list(dims)
        self.t = 0.0

    # ------------------------------------------------------------------ basics
    def _chk(self, rc):
        _lib.check(rc, self.h)

    def close(self):
        if getattr(self, "h", None):
            if self._owned:
                self._L.grhd_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def pshape(self):
        ng = self.ng
        return (self.nz + 2 * ng + 1, self.ny + 2 * ng + 1, self.nx + 2 * ng + 1, 5)

    @property
    def ushape(self):
        return (self.nz, self.ny, self.nx, 5)

    @property
    def n_cells(self) -> int:
        return self.nx * self.ny * self.nz

    # -------------------------------------------------------------- state I/O
    def initial_conditions(self):
        """set_initial_conditions + prim_to_cons (grhd2.f90:121-134)."""
        self._chk(self._L.grhd_initial_conditions(self.h))

    def perturb(self, kind: int, amplitude: float, seed: int):
        """Seeded noise: kind 1 pressure, 2 isentropic density (multiplicative),
        3 additive v^z (Kelvin-Helmholtz random seed)."""
        self._chk(self._L.grhd_perturb(self.h, int(kind), float(amplitude), int(seed)))

    def set_primitives(self, p: np.ndarray):
        self._chk(self._L.grhd_set_primitives(self.h, _ptr(np.ascontiguousarray(p, np.float64))))

    def put_primitives(self, p: np.ndarray):
        self._chk(self._L.grhd_put_primitives(self.h, _ptr(np.ascontiguousarray(p, np.float64))))

    def set_conserved(self, u: np.ndarray):
        self._chk(self._L.grhd_set_conserved(self.h, _ptr(np.ascontiguousarray(u, np.float64))))

    def put_conserved(self, u: np.ndarray):
        self._chk(self._L.grhd_put_conserved(self.h, _ptr(np.ascontiguousarray(u, np.float64))))

    def get_primitives(self) -> np.ndarray:
        p = np.zeros(self.pshape)
        self._chk(self._L.grhd_get_primitives(self.h, _ptr(p)))
        return p

    def get_conserved(self) -> np.ndarray:
        u = np.zeros(self.ushape)
        self._chk(self._L.grhd_get_conserved(self.h, _ptr(u)))
        return u

    # --------------------------------------------------------------- update
    def compute_dt(self) -> float:
        """update_adaptive_dt (evolution.f90:198-275)."""
        d = ctypes.c_double()
        self._chk(self._L.grhd_compute_dt(self.h, ctypes.byref(d)))
        return d.value

    def step(self, dt: float, nsteps: int = 1):
        self._chk(self._L.grhd_step(self.h, float(dt), int(nsteps)))

    def advance(self, t_final: float | None = None, max_steps: int = 2 ** 31 - 1):
        """The reference time loop (grhd2.f90:179-265) from self.t."""
        t_final = self.cfg.final_time if t_final is None else t_final
        t = ctypes.c_double(self.t)
        n = ctypes.c_int32()
        last = ctypes.c_double()
        self._chk(self._L.grhd_advance(self.h, float(t_final), int(max_steps), ctypes.byref(t),
                                       ctypes.byref(n), ctypes.byref(last)))
        self.t = t.value
        return n.value, last.value

    def boundaries(self):
        self._chk(self._L.grhd_boundaries(self.h))

    def rhs(self) -> np.ndarray:
        r = np.zeros(self.ushape)
        self._chk(self._L.grhd_rhs(self.h, _ptr(r)))
        return r

    def con2prim(self):
        self._chk(self._L.grhd_con2prim(self.h))

    def counters(self) -> dict:
        c = (ctypes.c_int64 * 4)()
        self._chk(self._L.grhd_counters(self.h, c))
        return {"atmosphere_resets": c[0], "brent_fallbacks": c[1], "steps": c[2]}

    def total_mass(self) -> float:
        """Mass monitor (grhd2.f90:283-301), reduced on the device."""
        m = ctypes.c_double()
        self._chk(self._L.grhd_total_mass(self.h, ctypes.byref(m)))
        return m.value

    def convergence_norms(self, t: float = 0.0):
        """calc_convergence_norms (output.f90:1144-1258): (L1, L2, Linf) of rho
        against the analytic solution at time t, and the Linf cell."""
        n = (ctypes.c_double * 3)()
        im = ctypes.c_int32()
        self._chk(self._L.grhd_convergence_norms(self.h, float(t), n, ctypes.byref(im)))
        return (n[0], n[1], n[2]), im.value

    def global_diagnostics(self) -> dict:
        """calc_global_diagnostics (output.f90:1052-1125), reduced on the device."""
        o = (ctypes.c_double * 8)()
        n = (ctypes.c_int64 * 2)()
        self._chk(self._L.grhd_global_diagnostics(self.h, o, n))
        keys = ("mass", "disk_mass", "angular_momentum", "rho_min", "p_min", "rho_max", "p_max", "v2_max")
        d = {k: o[i] for i, k in enumerate(keys)}
        d.update(atmosphere_cells=n[0], invalid_cells=n[1])
        return d

    def ppi_modes(self) -> np.ndarray:
        """calc_ppi_modes (output.f90:995-1049): c0, then per m = 1..4 Re, Im,
        amplitude and phase."""
        o = (ctypes.c_double * 17)()
        self._chk(self._L.grhd_ppi_modes(self.h, o))
        return np.array(o[:])

    def accretion_rate(self):
        """calc_m_dot (output.f90:951-993): (mdot, probe cell)."""
        m = ctypes.c_double()
        ip = ctypes.c_int32()
        self._chk(self._L.grhd_accretion_rate(self.h, ctypes.byref(m), ctypes.byref(ip)))
        return m.value, ip.value

    def synchronize(self):
        self._chk(self._L.grhd_synchronize(self.h))

    def set_stage_form(self, zseg: int):
        """-1: timed choice on the first step; 0: split form; n > 0: z-march
        form with n-cell segments; -2: x-fused form (3D grids above 65,536
        cells)."""
        self._chk(self._L.grhd_set_stage_form(self.h, int(zseg)))

    def set_face_form(self, form: int):
        """Face-kernel form (grhd_set_face_form): -1 timed, 1/2/4/8 LDS
        geometry groups, 0 geometry table, 16 table with deferred faces,
        32/48 forms 0/16 with the x and y faces walking z first, 64/80 with
        them walking bands of plane tiles."""
        self._chk(self._L.grhd_set_face_form(self.h, int(form)))

    def last_timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._chk(self._L.grhd_last_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value
