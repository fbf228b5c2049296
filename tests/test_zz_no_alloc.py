"""grhd_step_noalloc (no reference counterpart; for callers that cannot
allocate device memory) against grhd_step, through the C ABI.

Two calls in a row on one device share its scratch array: first a 3D
Kerr-Schild torus (MP5 + HLLC + sensor, FP64, 48 x 32 x 16, two steps), then a
smaller problem, Michel accretion on Kerr-Schild (96 cells, mixed precision,
three steps: the odd count ends in the stage buffer and is copied back, and
the outer injector comes from p's ghost layers).  Each must equal the same
steps through a context (grhd_step) bit for bit.  A problem whose
intermediates exceed the array is refused with GRHD_ERR_CAPACITY.  The
caller's device buffers come from hipMalloc of the HIP runtime the library
links (test plumbing; the entry point itself allocates nothing).
"""
import ctypes

import numpy as np
import pytest

from grhd_amd import _lib
from grhd_amd.params import config_from_par
from tests.cases import CASES, NOISY
from tests.helpers import read_set_mask, seeded_noise

GRHD_ERR_CAPACITY = 7


def _to_device_p(p, nx, ny, nz, ng):
    """Reference layout [NZg][NYg][NXg][5] -> device SoA [5][NZ][NY][NX]
    (an inactive axis keeps its single cell, reference index 1)."""
    jy = slice(None) if ny > 1 else slice(ng + 1, ng + 2)
    kz = slice(None) if nz > 1 else slice(ng + 1, ng + 2)
    return np.ascontiguousarray(p[kz, jy].transpose(3, 0, 1, 2))


def _from_device_p(pd, like, nx, ny, nz, ng):
    out = like.copy()
    jy = slice(None) if ny > 1 else slice(ng + 1, ng + 2)
    kz = slice(None) if nz > 1 else slice(ng + 1, ng + 2)
    out[kz, jy] = pd.transpose(1, 2, 3, 0)
    return out


def _case(name, extra=""):
    """The case's .par text with the keys of `extra` replaced or added."""
    keys = {ln.split("=")[0].strip(): ln for ln in extra.splitlines() if "=" in ln}
    lines = [ln for ln in CASES[name].splitlines() if ln.split("=")[0].strip() not in keys]
    return config_from_par("\n".join(lines + list(keys.values())) + "\n")


def _with_context(cfg, name, nsteps):
    from grhd_amd.solver import Solver
    with Solver(cfg) as S:
        S.initial_conditions()
        p0 = S.get_primitives()
        if name in NOISY:
            kind, amp, seed = NOISY[name]
            p0 = seeded_noise(p0, S.nx, S.ny, S.nz, S.ng, kind, amp, seed)
            S.set_primitives(p0)
        u0 = S.get_conserved()
        dt = S.compute_dt()
        S.step(dt, nsteps)
        return p0, u0, dt, S.get_conserved(), S.get_primitives(), S.ng


def _hip():
    """The HIP runtime libgrhd_hip.so links (already loaded with it)."""
    _lib.lib()
    with open("/proc/self/maps") as fh:
        path = next(ln.split()[-1] for ln in fh if "libamdhip64.so" in ln and "/torch/" not in ln)
    h = ctypes.CDLL(path)
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


class _DeviceArray:
    """A device copy of a float64 host array (hipMalloc / hipMemcpy)."""

    def __init__(self, hip, host):
        self.hip, self.shape = hip, host.shape
        self.nbytes = host.nbytes
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0
        h = np.ascontiguousarray(host, np.float64)
        assert hip.hipMemcpy(self.ptr, h.ctypes.data, self.nbytes, 1) == 0  # host to device

    def get(self):
        h = np.empty(self.shape, np.float64)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(h.ctypes.data, self.ptr, self.nbytes, 2) == 0  # device to host
        return h

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc(cfg, p0, u0, dt, nsteps, ng):
    hip = _hip()
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    pd = _DeviceArray(hip, _to_device_p(p0, nx, ny, nz, ng))
    ud = _DeviceArray(hip, u0.transpose(3, 0, 1, 2))
    try:
        c = _lib.Config.from_run(cfg)
        rc = _lib.lib().grhd_step_noalloc(ctypes.byref(c), pd.ptr, ud.ptr, dt, nsteps, None)
        assert rc == 0, _lib.lib().grhd_last_error(None)
        u = np.ascontiguousarray(ud.get().transpose(1, 2, 3, 0))
        p = _from_device_p(pd.get(), p0, nx, ny, nz, ng)
    finally:
        pd.free()
        ud.free()
    return u, p


@pytest.mark.gpu
def test_noalloc_step_equals_context_step_twice():
    runs = [("fm3d_ks_mp5_hllc_sensor", "nx = 48\nny = 32\nnz = 16\n", 2),
            ("michel_ks_a090", "precision = mixed\n", 3)]
    for name, extra, nsteps in runs:
        cfg = _case(name, extra)
        print(f"no-alloc step: {name} {cfg.nx}x{cfg.ny}x{cfg.nz} precision {cfg.precision}, "
              f"{nsteps} steps", flush=True)
        p0, u0, dt, u_ctx, p_ctx, ng = _with_context(cfg, name, nsteps)
        u, p = _noalloc(cfg, p0, u0, dt, nsteps, ng)
        assert np.array_equal(u, u_ctx), name
        m = read_set_mask(cfg.nx, cfg.ny, cfg.nz, ng)
        assert np.array_equal(p[m], p_ctx[m]), name


@pytest.mark.gpu
def test_noalloc_step_refuses_beyond_the_array():
    """Refused on the host before any device work (the buffers are never
    touched, so an 8-element one stands in for both)."""
    hip = _hip()
    cfg = _case("fm3d_ks_mp5_hllc_sensor", "nx = 256\nny = 128\nnz = 128\n")
    x = _DeviceArray(hip, np.zeros(8))
    c = _lib.Config.from_run(cfg)
    rc = _lib.lib().grhd_step_noalloc(ctypes.byref(c), x.ptr, x.ptr, 1e-3, 1, None)
    x.free()
    assert rc == GRHD_ERR_CAPACITY
    assert b"scratch array holds" in _lib.lib().grhd_last_error(None)
