#!/usr/bin/env python
"""Throughput of the SSP-RK3 update on the 3D Fishbone-Moncrief torus.

Workload (one "step" = one full SSP-RK3 step of every cell): Kerr-Schild
a = 0.9, logarithmic radial grid 256 x 128 x 128 (r, theta, phi) on
r in [1.2, 40], theta in [0.05 pi, 0.95 pi], phi in [0, 2 pi); Gamma = 4/3,
r_in = 6, r_max = 12, polytropic K with rho_max = 1; MP5 reconstruction with the
shock-sensor cascade, HLLC with HLLE fallback; seeded 1e-2 density noise.
State resident in HBM before timing.  Reports FP64 (headline) and mixed
FP32-face / FP64-state throughput with the mixed run's density error against
the FP64 run, the HBM roofline fraction of the stage kernels from the
algorithmic 560 B per cell-update, and the reference CPU path (OpenMP Fortran
built from its sources) on the same grid for a bounded number of steps on the
host cores.  Then a served ensemble: a seeded Poisson stream of independent
problems over the serving workload's full ranges (shock tubes nx 200-6400,
Michel nr 256-4096 with spin in [0, 0.99), Kelvin-Helmholtz 256^2-1024^2),
continuously batched on the device, with p50/p99 time-to-solution and
aggregate cell-updates/s ("ensemble").

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--serve-requests N]
                [--serve-rates 16] [--serve-mix full|quick]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "cell-updates/s on 3D FM torus (FP64 vs mixed); p50/p99 ensemble latency"
B_ALG = 560.0          # algorithmic HBM bytes per cell-update (FP64 state)
HBM_PEAK = 8.0e12      # MI355X HBM3E, B/s

C4_PAR = """problem = fishbone_3d
metric = kerr_schild
geometry = spheroidal
logarithmic_r = true
spin = 0.9
nx = {nx}
ny = {ny}
nz = {nz}
r_min = 1.2
r_max = 40
theta_min = {tmin}
theta_max = {tmax}
phi_min = 0
phi_max = 2
fm_inner_radius = 6
fm_pressure_max_radius = 12
reconstruction = mp5
riemann_solver = hllc
shock_sensor = true
cfl = 0.4
final_time = 1000
"""


def c4_config(nx=256, ny=128, nz=128, open_theta=True):
    from grhd_amd import params
    text = C4_PAR.format(nx=nx, ny=ny, nz=nz, tmin=0.05 if open_theta else 0,
                         tmax=0.95 if open_theta else 1)
    entries = params.parse_par_lines(text)
    cfg = params.preset(12)
    for _, k, v in entries:
        if k != "problem":
            params._apply_key(cfg, k, v, 12)
    if open_theta:
        cfg.extra["_allow_open_theta"] = True
    params.validate(cfg)
    return cfg


def run_gpu(cfg, precision, steps, warmup):
    from grhd_amd.solver import Solver, fishbone_unit_density_k
    cfg = cfg.replace(precision=precision)
    cfg.fm_polytropic_constant = fishbone_unit_density_k(cfg)
    with Solver(cfg) as S:
        S.initial_conditions()
        S.perturb(2, 1.0e-2, 3435)
        dt = S.compute_dt()
        S.step(dt, warmup)
        S.synchronize()
        t0 = time.perf_counter()
        S.step(dt, steps)
        S.synchronize()
        t1 = time.perf_counter()
        ev_ms, launches = S.last_timing()
        counters = S.counters()
        ncell = S.n_cells
        # final state (after the timed steps) for the mixed-vs-FP64 comparison
        D = S.get_conserved()[..., 0].copy()
        g = S.ng
        rho = S.get_primitives()[g + 1:g + 1 + S.nz, g + 1:g + 1 + S.ny, g + 1:g + 1 + S.nx, 0].copy()
    wall = t1 - t0
    return {
        "D": D,
        "rho": rho,
        "cell_updates_per_s": ncell * steps / wall,
        "wall_s": wall,
        "event_ms": ev_ms,
        "launches": launches,
        "n_cells": ncell,
        "dt": dt,
        "K": cfg.fm_polytropic_constant,
        "counters": counters,
    }


def cpu_child(out_path, nx, ny, nz, seconds):
    """Reference (oracle/_ref) on a bounded sample; runs in its own process."""
    import numpy as np
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    import oracle
    cfg = c4_config(nx, ny, nz, open_theta=False)  # the reference requires theta = [0, pi]
    t_setup = time.perf_counter()
    R = oracle.Reference(cfg)
    t_setup = time.perf_counter() - t_setup
    p = R.get_prim()
    rng = np.random.default_rng(3435)
    g = R.ng
    sl = (slice(g + 1, g + 1 + nz), slice(g + 1, g + 1 + ny), slice(g + 1, g + 1 + nx))
    rho = p[sl + (0,)]
    f = np.where(rho > 0.05 * rho.max(), 1.0 + 1e-2 * (2 * rng.random(rho.shape) - 1), 1.0)
    p[sl + (0,)] *= f
    p[sl + (1,)] *= f ** (4.0 / 3.0)
    R.set_prim(p)
    R.set_time(0.0, 1.0e9)
    t0 = time.perf_counter()
    R.step(1)
    one = time.perf_counter() - t0
    n = max(1, int(seconds / max(one, 1e-6)))
    t0 = time.perf_counter()
    R.step(n)
    el = time.perf_counter() - t0
    with open(out_path, "w") as fh:
        json.dump({"steps": n, "seconds": el, "cells": nx * ny * nz, "setup_s": t_setup}, fh)


def _host_description():
    """CPU model and the reference build's compiler (for the baseline record)."""
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as fh:
            model = next(ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name"))
    except (OSError, StopIteration):
        pass
    compiler = "amdflang"
    try:
        out = subprocess.run(["amdflang", "--version"], capture_output=True, text=True, timeout=30).stdout
        compiler = out.splitlines()[0].strip() if out else compiler
    except (OSError, subprocess.SubprocessError):
        pass
    return model, compiler


def cpu_baseline(seconds=15.0, nx=256, ny=128, nz=128):
    """The reference (oracle/_ref: the unmodified Fortran sources compiled by
    oracle/Makefile with amdflang -O2 -ffp-contract=off -fopenmp) on the
    torus grid the GPU line measures, theta = [0, pi] (the reference refuses
    the open theta range), for as many RK3 steps as fit in `seconds` after one
    timed step, with OpenMP on the host cores this process may use."""
    import oracle
    if not oracle.reference_available():
        return None
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close")
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "cpu.json")
        cmd = [sys.executable, str(Path(__file__).resolve()), "--cpu-child", out,
               "--cpu-seconds", str(seconds), "--cpu-grid", f"{nx},{ny},{nz}"]
        subprocess.run(cmd, env=env, check=True, timeout=900, cwd=str(ROOT))
        with open(out) as fh:
            r = json.load(fh)
    model, compiler = _host_description()
    return {
        "value": r["cells"] * r["steps"] / r["seconds"],
        "unit": "cell-updates/s",
        "cores": cores,
        "kind": "reference",
        "cpu": model,
        "compiler": compiler,
        "flags": "-O2 -ffp-contract=off -fopenmp (OMP_PROC_BIND=close)",
        "sample": (f"reference Fortran, OpenMP {cores} threads, fishbone_3d KS a=0.9 {nx}x{ny}x{nz} "
                   f"theta=[0,pi], MP5+HLLC+sensor, 1e-2 density noise: {r['steps']} RK3 steps in "
                   f"{r['seconds']:.1f} s after one untimed step (set-up {r.get('setup_s', 0.0):.1f} s)"),
    }


# Served mixes.  "quick": Michel nr 256/512 and KH 128^2/256^2, so that the
# default run finishes in seconds; "full": the ranges of the serving workload
# (Michel nr 256-4096, KH 256^2-1024^2; the Michel nr = 4096 runs take about
# 945k steps each).
MIXES = {
    "quick": dict(shock_nx=(200, 400, 800, 1600, 3200, 6400), michel_nr=(256, 512), khi_n=(128, 256)),
    "full": dict(shock_nx=(200, 400, 800, 1600, 3200, 6400), michel_nr=(256, 512, 1024, 2048, 4096),
                 khi_n=(256, 512, 1024)),
}


def _progress(c):
    print(f"served {c.label}: {c.steps} steps, latency {c.latency:.2f} s", file=sys.stderr, flush=True)


def run_ensemble(n, rate, time_scale, chunk, seed=3435, mix="quick", progress=False, pacing=8):
    """Served request mix: Poisson arrivals of independent problems (shock
    tubes, Michel accretion with spin 0 or in [0, 0.99), Kelvin-Helmholtz with
    seeded v^z noise), continuously batched on the device; p50/p99
    time-to-solution and aggregate cell-updates/s."""
    from grhd_amd.batch import Batch, footprint
    from grhd_amd.serving import Server, request_mix
    m = MIXES[mix]
    reqs = request_mix(n, rate=rate, seed=seed, time_scale=time_scale, **m)
    with Batch(capacity=64) as B:
        B.set_pacing(pacing)
        rep = Server(B, footprint, hbm_budget=64 << 30, chunk=chunk,
                     on_complete=_progress if progress else None).run(reqs)
        info = B.info()
    s = rep.summary()
    lo, hi = min(m["shock_nx"]), max(m["shock_nx"])
    s.update({"arrival_rate_per_s": rate, "final_time_scale": time_scale, "chunk_steps": chunk,
              "pacing": pacing,
              "launch_form": info["variant"], "seed": seed,
              "mix": f"{mix}: sod/strong_shock/shu_osher nx {lo}-{hi}, michel nr "
                     f"{'/'.join(map(str, m['michel_nr']))} a=0 or U[0,0.99), kelvin_helmholtz "
                     f"{'/'.join(f'{k}^2' for k in m['khi_n'])} v=+-0.5 Gamma=4/3 with 1e-2 seeded v^z "
                     "noise, each to its preset final_time"})
    return s


def measured_traffic(precision="fp64"):
    """HBM bytes per RK3 stage from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (profiles/c4_hbm_counters.md, scaled by a plain copy of
    known size); None when the summary is absent."""
    import re
    path = ROOT / "profiles" / "c4_hbm_counters.md"
    try:
        text = path.read_text()
    except OSError:
        return None
    m = re.search(rf"TRAFFIC {precision} per RK3 stage: [^=]* written = ([0-9.]+) GB", text)
    return float(m.group(1)) * 1e9 if m else None


def michel_convergence(resolutions=(256, 1024, 4096)):
    """C2: Michel accretion (EF, log r in [1.8, 20], r_c = 8, WENO5 + HLLE) to
    t = 500 from the analytic state with 1e-3 seeded density noise, FP64 and
    mixed, at each resolution: L1/L2/Linf density error against the analytic
    solution (output.f90:1189-1213) beside the reference's where
    tests/golden holds it.  Each run is a batch of one problem (graph-replayed
    chunks; bit-identical to grhd_advance)."""
    import numpy as np
    from grhd_amd.batch import Batch
    from grhd_amd.params import config_from_par
    from tests.cases import MICHEL_C2
    from tests.helpers import michel_norms, seeded_noise
    out = []
    for nr in resolutions:
        gpath = ROOT / "tests" / "golden" / f"michel_c2_{nr}.npz"
        g = np.load(gpath) if gpath.exists() else None
        for prec in ("fp64", "mixed"):
            cfg = config_from_par(MICHEL_C2.format(nr=nr) + f"precision = {prec}\n")
            with Batch(capacity=1) as B:
                slot = B.add(cfg)
                view = B.problem(slot)
                exact = view.get_primitives()
                p0 = g["p0"] if g is not None else seeded_noise(exact, view.nx, view.ny, view.nz, view.ng, 2,
                                                                  1.0e-3, 3435)
                view.set_primitives(p0)
                t0 = time.perf_counter()
                while not B.status(slot)[0]:
                    B.step(4096)
                wall = time.perf_counter() - t0
                done, t, steps = B.status(slot)
                norms = michel_norms(view.get_primitives(), exact, cfg.nx, cfg.ny, view.ng)
            rec = {"nr": nr, "precision": prec, "steps": steps, "t": t, "wall_s": wall,
                   "L1": norms[0], "L2": norms[1], "Linf": norms[2],
                   "reference_L1": float(g["norms"][0]) if g is not None else None,
                   "reference_steps": int(g["n_steps"]) if g is not None else None}
            print(json.dumps(rec), flush=True)
            out.append(rec)
    return out


C3_PAR = """problem = fishbone_equatorial
fm_inner_radius = 6
fm_pressure_max_radius = 12
reconstruction = weno5
riemann_solver = hlle
"""


def equatorial_torus(steps=1000, batch_forms=True):
    """C3: Fishbone-Moncrief equatorial torus, Kerr-Schild a = 0.9, log-r
    spheroidal grid 400 x 1 x 200 (the reference's preset grid), r_in = 6,
    pressure maximum at r = 12, K for rho_max = 1, 1e-2 seeded pressure noise
    inside the torus, WENO5 + HLLE, `steps` RK3 steps in FP64 and in mixed
    precision: throughput, and the mixed run's conserved density D against the
    FP64 run's (relative L1).  The floors are the reference's constants (rho 1e-10,
    p 1e-13); r-dependent floors are not a feature of the reference."""
    import numpy as np
    from grhd_amd.batch import Batch
    from grhd_amd.params import config_from_par
    from grhd_amd.solver import Solver, fishbone_unit_density_k
    out, rho = {}, {}
    for prec in ("fp64", "mixed"):
        cfg = config_from_par(C3_PAR + f"precision = {prec}\n")
        cfg.fm_polytropic_constant = fishbone_unit_density_k(cfg)
        with Solver(cfg) as W:          # loads the kernels outside the timed run
            W.initial_conditions()
            W.step(W.compute_dt(), 3)
            W.synchronize()
        with Solver(cfg) as S:
            S.initial_conditions()
            S.perturb(1, 1.0e-2, 3435)
            dt = S.compute_dt()
            S.synchronize()
            t0 = time.perf_counter()
            S.step(dt, steps)
            S.synchronize()
            wall = time.perf_counter() - t0
            ev_ms, _ = S.last_timing()
            u_direct = S.get_conserved()
            rho[prec] = u_direct[..., 0]
            out[prec] = {"cell_updates_per_s": S.n_cells * steps / wall, "ms_per_step": 1e3 * wall / steps,
                         "event_ms_per_step": ev_ms / steps, "steps": steps, "t": steps * dt,
                         **S.counters()}
            ncell = S.n_cells
        # the same run as a batch of one: graph-replayed chunks of 64 steps in
        # each launch form (capture included in the time); the state must equal
        # the direct run's bit for bit
        forms = {}
        for v in (range(1, 6) if batch_forms else ()):
            with Batch(capacity=1) as B:
                B.set_variant(v)
                slot = B.add(cfg)
                view = B.problem(slot)
                view.perturb(1, 1.0e-2, 3435)
                view.synchronize()
                t0 = time.perf_counter()
                B.step(steps)
                taken = B.status(slot)[2]
                wall_b = time.perf_counter() - t0
                assert taken == steps and np.array_equal(view.get_conserved(), u_direct), (prec, v)
            forms[v] = ncell * steps / wall_b
        if forms:
            best = max(forms, key=forms.get)
            out[prec]["batched_cell_updates_per_s"] = forms[best]
            out[prec]["batched_launch_form"] = best
            out[prec]["batched_by_form"] = {str(k): round(x / 1e6, 1) for k, x in forms.items()}
    a, b = rho["mixed"], rho["fp64"]
    rec = {"workload": "fishbone_equatorial Kerr-Schild a=0.9 log-r 400x1x200, r_in=6, r_pmax=12, "
                       "K for rho_max=1, 1e-2 pressure noise, WENO5+HLLE",
           **out, "mixed_speedup": out["mixed"]["cell_updates_per_s"] / out["fp64"]["cell_updates_per_s"],
           "mixed_vs_fp64_D_L1": float(np.abs(a - b).sum() / np.abs(b).sum())}
    print(json.dumps(rec), flush=True)
    return rec


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain device-to-device copy of nbytes, reps times: its FETCH_SIZE and
    WRITE_SIZE per dispatch give the counters' scale against known traffic
    (nbytes read and nbytes written per dispatch)."""
    import torch
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", default="256,128,128")
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-grid", default="256,128,128")
    ap.add_argument("--cpu-child", default=None)
    ap.add_argument("--serve-requests", type=int, default=32,
                    help="requests of the served ensemble (0: skip it)")
    ap.add_argument("--serve-rates", default="16",
                    help="Poisson arrival rates (requests/s) of the served ensemble, swept in order; "
                         "the line's 'ensemble' is the one at 16/s (or the first)")
    ap.add_argument("--serve-mix", choices=sorted(MIXES), default="full")
    ap.add_argument("--serve-pacing", type=int, default=None,
                    help="steps of a cheaper batch group per step of the most expensive one (1..8); "
                         "default 8 for the full mix, whose problem sizes span 200 to 1M cells, "
                         "and 1 for the quick mix, where pacing raised the median time-to-solution")
    ap.add_argument("--serve-progress", action="store_true",
                    help="print each served request to stderr as it completes")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the served ensemble (no torus throughput, no CPU baseline)")
    ap.add_argument("--serve-time-scale", type=float, default=1.0)
    ap.add_argument("--serve-chunk", type=int, default=32)
    ap.add_argument("--serve-seed", type=int, default=3435,
                    help="seed of the request stream (arrivals and request draws)")
    ap.add_argument("--c3", action="store_true",
                    help="only the equatorial torus: 1000 RK3 steps in FP64 and mixed precision")
    ap.add_argument("--c3-direct", action="store_true",
                    help="with --c3: the direct path only (no batch-of-one launch forms)")
    ap.add_argument("--michel", action="store_true",
                    help="only the Michel C2 convergence study (nr 256/1024/4096, FP64 and mixed)")
    ap.add_argument("--valu-rate", action="store_true",
                    help="only run the VALU issue-rate probe (grhd_amd/bin/valu_rate): cycles per wave64 "
                         "FP64/FP32 instruction per SIMD, the ceiling of the VALU-bound stage kernels")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (for HBM counter scaling)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.valu_rate:
        probe = ROOT / "grhd_amd" / "bin" / "valu_rate"
        if not probe.exists():
            raise SystemExit(f"{probe} is not built: run `python -m grhd_amd.build`")
        raise SystemExit(subprocess.call([str(probe)]))
    if a.michel:
        michel_convergence()
        return
    if a.c3:
        equatorial_torus(batch_forms=not a.c3_direct)
        return
    cnx, cny, cnz = (int(v) for v in a.cpu_grid.split(","))
    if a.cpu_child:
        cpu_child(a.cpu_child, cnx, cny, cnz, a.cpu_seconds)
        return
    if a.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {a.gpus}); "
              "sharding by radial slabs across GPUs is not implemented", file=sys.stderr)
        sys.exit(2)
    if a.serve_pacing is None:
        a.serve_pacing = 8 if a.serve_mix == "full" else 1
    if a.serve_only:
        for r in (float(r) for r in a.serve_rates.split(",") if r):
            print(json.dumps(run_ensemble(a.serve_requests, r, a.serve_time_scale, a.serve_chunk,
                                          seed=a.serve_seed, mix=a.serve_mix,
                                          progress=a.serve_progress, pacing=a.serve_pacing)),
                  flush=True)
        return
    nx, ny, nz = (int(v) for v in a.grid.split(","))
    cfg = c4_config(nx, ny, nz)
    res = {}
    if a.precision in ("fp64", "both"):
        res["fp64"] = run_gpu(cfg, 0, a.steps, a.warmup)
    if a.precision in ("mixed", "both"):
        res["mixed"] = run_gpu(cfg, 1, a.steps, a.warmup)
    head = res.get("fp64") or res["mixed"]
    stage_ms = head["event_ms"] / (3 * a.steps)
    achieved = B_ALG / 3.0 * head["n_cells"] / (stage_ms * 1e-3)
    line = {
        "metric": METRIC,
        "value": head["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["wall_s"] * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if "fp64" in res else "f32-face/f64-state",
        "data": "synthetic",
        "config": {
            "workload": f"fishbone_3d Kerr-Schild a=0.9 log-r {nx}x{ny}x{nz} theta=[0.05pi,0.95pi] "
                        "MP5+HLLC(HLLE fallback)+shock sensor, 1e-2 density noise, SSP-RK3",
            "n_cells": head["n_cells"],
            "precision": "fp64" if "fp64" in res else "mixed",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": measured_traffic("fp64" if "fp64" in res else "mixed"),
            "traffic_unit": "bytes per RK3 stage (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "kernel": "RK3 stage (face fluxes + cell update + con2prim + ghost fill)",
            "algorithmic_bytes": B_ALG / 3.0 * head["n_cells"],
            "stage_ms": stage_ms,
        },
    }
    if "mixed" in res:
        line["mixed"] = {
            "value": res["mixed"]["cell_updates_per_s"],
            "ms_per_step": res["mixed"]["wall_s"] * 1e3 / a.steps,
            "speedup_vs_fp64": (res["mixed"]["cell_updates_per_s"] / res["fp64"]["cell_updates_per_s"]
                                if "fp64" in res else None),
        }
        if "fp64" in res:
            import numpy as np
            # both runs start from the same seeded state and take the same
            # warmup + timed steps at the same dt
            for key, name in (("D", "D_L1_vs_fp64"), ("rho", "rho_L1_vs_fp64")):
                a64, amx = res["fp64"][key], res["mixed"][key]
                line["mixed"][name] = float(np.abs(amx - a64).sum() / np.abs(a64).sum())
            line["mixed"]["compared_after_steps"] = a.warmup + a.steps
    if a.serve_requests > 0:
        rates = [float(r) for r in a.serve_rates.split(",") if r]
        sweep = [run_ensemble(a.serve_requests, r, a.serve_time_scale, a.serve_chunk, seed=a.serve_seed,
                              mix=a.serve_mix, progress=a.serve_progress, pacing=a.serve_pacing)
                 for r in rates]
        line["ensemble"] = next((e for e in sweep if e["arrival_rate_per_s"] == 16.0), sweep[0])
        line["ensemble_sweep"] = [{k: e[k] for k in ("arrival_rate_per_s", "p50_s", "p99_s", "mean_wait_s",
                                                     "cell_updates_per_s", "wall_s", "peak_active")}
                                  for e in sweep]
    line["cpu_baseline"] = None if a.no_cpu_baseline else cpu_baseline(a.cpu_seconds, cnx, cny, cnz)
    line["counters"] = head["counters"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
