"""Continuous-batching request loop over one device.

A request is one problem (a ``RunConfig``) that arrives at some time and
wants its final state.  :class:`Server` keeps a :class:`~grhd_amd.batch.Batch`
of resident problems and loops:

1. requests whose arrival time has passed join the queue;
2. queued requests are admitted first-fit in arrival order while a slot is
   free and their HBM footprint fits the remaining budget -- requests that
   arrive while others run join the running batch at the next chunk boundary;
3. the batch advances every unfinished problem by ``chunk`` steps (their work
   shares the same launches on the device);
4. finished problems are read back, released, and their time-to-solution
   (completion - arrival) is recorded.

The result reports p50/p99 time-to-solution beside aggregate cell-updates/s.
The scheduling logic takes any object with the Batch methods used here, so it
is tested on the CPU with a stand-in.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import params as P
from .params import RunConfig


@dataclass
class Request:
    rid: int
    cfg: RunConfig
    arrival: float                      # seconds after the loop starts
    label: str = ""
    perturb: tuple | None = None        # (kind, amplitude, seed) injected at admission
    want_state: bool = False            # keep the final conserved state


@dataclass
class Completed:
    rid: int
    label: str
    arrival: float
    admitted: float
    finished: float
    steps: int
    cells: int
    state: np.ndarray | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def wait(self) -> float:
        return self.admitted - self.arrival


@dataclass
class Report:
    completed: list = field(default_factory=list)
    wall: float = 0.0
    chunks: int = 0
    peak_active: int = 0
    peak_bytes: int = 0

    def latencies(self) -> np.ndarray:
        return np.array([c.latency for c in self.completed])

    def summary(self) -> dict:
        lat = self.latencies()
        cu = sum(c.cells * c.steps for c in self.completed)
        return {
            "requests": len(self.completed),
            "p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "mean_wait_s": float(np.mean([c.wait for c in self.completed])) if len(lat) else None,
            "cell_updates_per_s": cu / self.wall if self.wall > 0 else None,
            "wall_s": self.wall,
            "chunks": self.chunks,
            "peak_active": self.peak_active,
            "peak_hbm_bytes": self.peak_bytes,
        }


class Server:
    def __init__(self, batch, footprint, hbm_budget: int, chunk: int = 32, clock=time.perf_counter,
                 sleep=time.sleep, on_complete=None):
        self.batch = batch
        self.on_complete = on_complete     # called with each Completed record
        self.footprint = footprint
        self.budget = int(hbm_budget)
        self.chunk = int(chunk)
        self.clock = clock
        self.sleep = sleep

    def run(self, requests) -> Report:
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in pending:
            if self.footprint(r.cfg) > self.budget:
                raise ValueError(f"request {r.rid} needs more than the whole HBM budget")
        queue: list[Request] = []
        active: dict[int, tuple[Request, float, int]] = {}   # slot -> (request, admitted, bytes)
        used = 0
        rep = Report()
        t0 = self.clock()
        i = 0
        while i < len(pending) or queue or active:
            now = self.clock() - t0
            while i < len(pending) and pending[i].arrival <= now:
                queue.append(pending[i])
                i += 1
            # first-fit admission in arrival order
            still = []
            for r in queue:
                need = self.footprint(r.cfg)
                if len(active) < self.batch.capacity and used + need <= self.budget:
                    slot = self.batch.add(r.cfg)
                    if r.perturb is not None:
                        self.batch.problem(slot).perturb(*r.perturb)
                    active[slot] = (r, self.clock() - t0, need)
                    used += need
                else:
                    still.append(r)
            queue = still
            rep.peak_active = max(rep.peak_active, len(active))
            rep.peak_bytes = max(rep.peak_bytes, used)
            if not active:
                if i < len(pending):  # idle until the next arrival
                    self.sleep(max(0.0, pending[i].arrival - (self.clock() - t0)))
                continue
            self.batch.step(self.chunk)
            rep.chunks += 1
            for slot in list(active):
                done, _, steps = self.batch.status(slot)
                if not done:
                    continue
                r, admitted, need = active.pop(slot)
                state = None
                view = self.batch.problem(slot)
                if r.want_state:
                    state = view.get_conserved()
                cells = view.n_cells
                self.batch.remove(slot)
                used -= need
                rep.completed.append(Completed(r.rid, r.label, r.arrival, admitted, self.clock() - t0, steps,
                                               cells, state))
                if self.on_complete is not None:
                    self.on_complete(rep.completed[-1])
        rep.wall = self.clock() - t0
        return rep


# ------------------------------------------------------------ request mix --
def _shock_tube(problem: int, nx: int) -> RunConfig:
    cfg = P.preset(problem).replace(nx=nx)
    P.validate(cfg)
    return cfg


def _michel(nr: int, spin: float) -> RunConfig:
    cfg = P.preset(7).replace(nx=nr, ny=1, nz=1, log_r=True, r_min=1.8, r_max=20.0,
                              michel_critical_radius=8.0)
    if spin > 0.0:
        cfg = cfg.replace(metric=P.KERR_SCHILD, geometry=P.SPHEROIDAL, spin=spin)
    P.validate(cfg)
    return cfg


def _kelvin_helmholtz(n: int) -> RunConfig:
    cfg = P.preset(4).replace(nx=n, nz=n, adiabatic_index=4.0 / 3.0, khi_velocity_inner=0.5,
                              khi_velocity_outer=-0.5)
    P.validate(cfg)
    return cfg


def request_mix(n: int, rate: float, seed: int = 3435, time_scale: float = 1.0,
                shock_nx=(200, 400, 800, 1600, 3200, 6400), michel_nr=(256, 512),
                khi_n=(128, 256), khi_noise: float = 1.0e-2) -> list[Request]:
    """n requests with Poisson arrivals (rate per second), drawn uniformly over
    three kinds: a shock tube (sod, strong_shock or shu_osher) with nx from
    shock_nx; Michel accretion with nr from michel_nr and spin in [0, 0.99)
    (Kerr-Schild spheroidal when a > 0); Kelvin-Helmholtz with N x N cells
    from khi_n, v = +-0.5, Gamma = 4/3, whose sinusoidal seed perturbation
    gets seeded white noise of amplitude khi_noise on v^z at admission.  Each
    runs to its preset final time times time_scale."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += float(rng.exponential(1.0 / rate))
        kind = int(rng.integers(3))
        if kind == 0:
            prob = int(rng.choice([1, 2, 3]))
            nx = int(rng.choice(shock_nx))
            cfg = _shock_tube(prob, nx)
            label = f"{['sod', 'strong_shock', 'shu_osher'][prob - 1]} nx={nx}"
        elif kind == 1:
            nr = int(rng.choice(michel_nr))
            spin = float(rng.uniform(0.0, 0.99)) if rng.random() < 0.5 else 0.0
            cfg = _michel(nr, spin)
            label = f"michel nr={nr} a={spin:.2f}"
        else:
            nk = int(rng.choice(khi_n))
            cfg = _kelvin_helmholtz(nk)
            label = f"kelvin_helmholtz {nk}x{nk}"
        if time_scale != 1.0:
            cfg = cfg.replace(final_time=cfg.final_time * time_scale)
        noise = (3, khi_noise, seed * 1000003 + rid) if kind == 2 and khi_noise > 0.0 else None
        out.append(Request(rid, cfg, t, label, perturb=noise))
    return out


def percentile(x, q) -> float:
    return float(np.percentile(np.asarray(x, float), q)) if len(x) else math.nan
