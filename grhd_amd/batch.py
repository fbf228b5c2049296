"""Many independent problems on one device (C ABI ``grhd_batch_*``).

:class:`Batch` keeps problems of any kind and grid size resident and advances
all unfinished ones together, each with its own dt and time: the RK3 stages of
every problem of one metric kind go into shared launches (hipGraph-captured;
the launch form is autotuned).  A problem's results are bit-identical to
running it alone through :meth:`Solver.advance`.
"""
from __future__ import annotations

import ctypes

from . import _lib
from .params import RunConfig
from .solver import Solver

GRHD_ERR_CAPACITY = 7
VARIANTS = {0: "autotuned", 1: "phase kernels (6 launches per stage)",
            2: "fused phase kernels (3 launches per stage), 4 waves per geometry column",
            3: "fused phase kernels (3 launches per stage), 1 wave per geometry column",
            4: "persistent stage kernel, 4 waves per geometry column",
            5: "persistent stage kernel, 1 wave per geometry column"}


def footprint(cfg: RunConfig) -> int:
    """Device bytes a problem's state takes (grhd_footprint)."""
    c = _lib.Config.from_run(cfg)
    n = ctypes.c_uint64()
    _lib.check(_lib.lib().grhd_footprint(ctypes.byref(c), ctypes.byref(n)))
    return n.value


class CapacityError(_lib.GrhdError):
    """No free slot, or the problem does not fit the remaining HBM budget."""


class Batch:
    def __init__(self, capacity: int = 64, hbm_budget: int = 0):
        self._L = _lib.lib()
        h = ctypes.c_void_p()
        rc = self._L.grhd_batch_create(int(capacity), int(hbm_budget), ctypes.byref(h))
        if rc != _lib.GRHD_OK:
            raise _lib.GrhdError(rc, "grhd_batch_create failed")
        self.h = h
        self.capacity = capacity
        self._cfg = {}

    def _chk(self, rc):
        if rc != _lib.GRHD_OK:
            msg = self._L.grhd_batch_last_error(self.h)
            msg = msg.decode() if msg else ""
            if rc == GRHD_ERR_CAPACITY:
                raise CapacityError(rc, msg)
            raise _lib.GrhdError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            self._L.grhd_batch_destroy(self.h)
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

    def add(self, cfg: RunConfig) -> int:
        """Admits a problem with its initial conditions; returns its slot."""
        c = _lib.Config.from_run(cfg)
        slot = ctypes.c_int32()
        self._chk(self._L.grhd_batch_add(self.h, ctypes.byref(c), ctypes.byref(slot)))
        self._cfg[slot.value] = cfg
        return slot.value

    def problem(self, slot: int) -> Solver:
        """A view of the problem's context for state I/O (set/get, perturb)."""
        h = ctypes.c_void_p()
        self._chk(self._L.grhd_batch_problem(self.h, int(slot), ctypes.byref(h)))
        return Solver(self._cfg[slot], handle=h)

    def remove(self, slot: int):
        self._chk(self._L.grhd_batch_remove(self.h, int(slot)))
        self._cfg.pop(slot, None)

    def step(self, nsteps: int):
        self._chk(self._L.grhd_batch_step(self.h, int(nsteps)))

    def status(self, slot: int):
        """(done, t, steps) of a problem."""
        d = ctypes.c_int32()
        t = ctypes.c_double()
        n = ctypes.c_int64()
        self._chk(self._L.grhd_batch_status(self.h, int(slot), ctypes.byref(d), ctypes.byref(t),
                                            ctypes.byref(n)))
        return bool(d.value), t.value, n.value

    def set_variant(self, variant: int):
        self._chk(self._L.grhd_batch_set_variant(self.h, int(variant)))

    def set_pacing(self, max_ratio: int):
        """Cheaper groups take up to max_ratio steps per step of the most
        expensive group (1: every problem advances the same steps)."""
        self._chk(self._L.grhd_batch_set_pacing(self.h, int(max_ratio)))

    def info(self) -> dict:
        v = ctypes.c_int32()
        a = ctypes.c_int32()
        used = ctypes.c_uint64()
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._chk(self._L.grhd_batch_info(self.h, ctypes.byref(v), ctypes.byref(a), ctypes.byref(used),
                                          ctypes.byref(ms), ctypes.byref(n)))
        return {"variant": v.value, "active": a.value, "hbm_used": used.value, "last_ms": ms.value,
                "launches": n.value}
