"""Write the golden vectors under tests/golden/ from the reference build.

Runs the unmodified reference (oracle/_ref/libgrhd_ref.so, built from the
reference sources by oracle/Makefile) on the parity cases and stores inputs
and outputs as compressed .npz fixtures:

  <case>.npz : p0 (ghosted primitives fed in), u1 (conserved after one RK3
               step), p1, u10 (after ten steps), dt, t10
  sod_c1_final.npz : C1 to t = 0.4 (u, p row, number of steps)
  michel_c2_<nr>.npz: C2 at nr = 256 (and 1024) to t = 500 with seeded 1e-3
               density noise: p0, final u, the reference's L1/L2/Linf density
               error against the analytic solution (output.f90:1189-1213)
  con2prim_<case>.npz: cons_to_prim over every interior cell of a case's
               initial conserved state with extreme and invalid states planted
               in some cells: ultra-hot states (tau/D >= 1e24, S up to 1.5x
               its bound) whose rtsafe solve does not converge in 100
               iterations and falls back to Brent (equations.f90:337-339,
               343-424), and NaN, infinite, negative-D and negative-tau states
               that reset to the atmosphere (equations.f90:116-124, 186-204):
               u_in, u_out (repaired in place), p_out (interior), cells
  dust_<nr>.npz: stationary dust accretion (problem 8, EF, r in [1, 51]) at
               nr = 128, 256, 512 to t = 500: final u, p row and the dust
               norms against the analytic solution (output.f90:1214-1235)

Usage: python tests/golden/make_golden.py [case ...]   (OMP_NUM_THREADS=1)
"""
import os
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from grhd_amd.params import config_from_par  # noqa: E402
from tests.cases import CASES, DUST, MICHEL_C2, NOISY  # noqa: E402
from tests.helpers import dust_exact_density, dust_norms  # noqa: E402
from tests.helpers import michel_norms as helpers_michel_norms  # noqa: E402
from tests.helpers import seeded_noise  # noqa: E402

OUT = Path(__file__).resolve().parent
STEP_CASES = ["sod_c1", "fm3d_ks_mp5_hllc_sensor", "michel_ef_log", "khi_weno3_sensor",
              "jet_cylindrical", "fm_sagittal_ef_log", "shu_tvd_minmod_sensor"]


def michel_norms(p, cfg, ng):
    """Norms of p against the reference's analytic initial data."""
    R = oracle.Reference(cfg)
    try:
        exact = R.get_prim()
    finally:
        R.close()
    return helpers_michel_norms(p, exact, cfg.nx, cfg.ny, ng)


def step_case(name):
    cfg = config_from_par(CASES[name])
    R = oracle.Reference(cfg)
    try:
        p0 = R.get_prim()
        if name in NOISY:
            kind, amp, seed = NOISY[name]
            p0 = seeded_noise(p0, R.nx, R.ny, R.nz, R.ng, kind, amp, seed)
            R.set_prim(p0)
        t_final = cfg.final_time if cfg.final_time > 0 else 1.0
        R.set_time(0.0, t_final)
        dt = R.dt()
        R.step(1)
        u1, p1 = R.get_cons(), R.get_prim()
        t10, _, _ = R.step(9)
        u10 = R.get_cons()
    finally:
        R.close()
    np.savez_compressed(OUT / f"{name}.npz", p0=p0, u1=u1, p1=p1, u10=u10, dt=dt, t10=t10)


def sod_final():
    cfg = config_from_par(CASES["sod_c1"])
    R = oracle.Reference(cfg)
    try:
        p0 = R.get_prim()
        R.set_time(0.0, cfg.final_time)
        t, dt, n = R.step(10 ** 6)
        u, p = R.get_cons(), R.get_prim()
    finally:
        R.close()
    g = R.ng
    np.savez_compressed(OUT / "sod_c1_final.npz", p0=p0, u=u, p_row=p[g + 1, g + 1], t=t,
                        n_steps=n, last_dt=dt)


def michel_c2(nr=256):
    cfg = config_from_par(MICHEL_C2.format(nr=nr))
    R = oracle.Reference(cfg)
    try:
        p0 = seeded_noise(R.get_prim(), R.nx, R.ny, R.nz, R.ng, 2, 1.0e-3, 3435)
        R.set_prim(p0)
        R.set_time(0.0, cfg.final_time)
        t, dt, n = R.step(10 ** 8)
        u, p = R.get_cons(), R.get_prim()
        ng = R.ng
    finally:
        R.close()
    norms = michel_norms(p, cfg, ng)
    np.savez_compressed(OUT / f"michel_c2_{nr}.npz", p0=p0, u=u, p_row=p[ng + 1, ng + 1], t=t,
                        n_steps=n, norms=norms)
    print(f"michel nr={nr}: {n} steps, L1/L2/Linf = {norms}")


CON2PRIM_CASES = ["fm3d_ks_mp5_hllc_sensor", "michel_ks_a090", "sod_mp5_hllc_sensor"]


def con2prim_states():
    """Conserved states planted by con2prim_fixture: (D, tau, S) with tau/D
    from 1e24 to 1e40 and |S| a fraction of its bound sqrt(E^2 - D^2)
    (beyond 1 exercises the momentum rescale), in three directions; then
    states the root solve cannot take (NaN, inf, D < 0, tau < 0)."""
    nan, inf = np.nan, np.inf
    out = [[nan, 1.0, 0.0, 0.0, 0.0], [1.0, nan, 0.1, 0.0, 0.0], [1.0, 1.0, nan, 0.0, 0.0],
           [-1.0, 1.0, 0.0, 0.0, 0.0], [1.0, -5.0, 0.1, 0.0, 0.0], [1.0, inf, 0.0, 0.0, 0.0],
           [inf, 1.0, 0.0, 0.0, 0.0], [1.0, 1.0, 0.0, inf, 0.0], [0.0, 0.0, 0.0, 0.0, 0.0]]
    dirs = [(1.0, 0.0, 0.0), (0.6, 0.8, 0.0), (0.3, 0.2, 0.93)]
    for D in (1.0e-6, 1.0e-3, 1.0, 1.0e3):
        for ratio in (1.0e24, 1.0e30, 1.0e40):
            tau = D * ratio
            smax = np.sqrt((tau + D) ** 2 - D * D)
            for frac in (0.5, 0.999, 1.0, 1.5):
                for d in dirs:
                    out.append([D, tau, *(frac * smax * np.array(d))])
    return np.array(out)


def con2prim_fixture(name):
    """The reference's cons_to_prim over the interior (ref_con2prim_all) on the
    case's initial conserved state with the first min(153, ncell / 2) of
    con2prim_states() planted at spread cells (flat interior index
    (37 m + 5) mod ncell)."""
    cfg = config_from_par(CASES[name])
    R = oracle.Reference(cfg)
    try:
        u = R.get_cons()
        flat = u.reshape(-1, 5)
        ncell = flat.shape[0]
        states = con2prim_states()[:ncell // 2]
        cells = (37 * np.arange(len(states)) + 5) % ncell
        assert len(set(cells.tolist())) == len(states)
        flat[cells] = states
        u_in = u.copy()
        R.put_cons(u)
        R.con2prim_all()
        u_out, p = R.get_cons(), R.get_prim()
        g = R.ng
        sl = (slice(g + 1, g + 1 + R.nz), slice(g + 1, g + 1 + R.ny), slice(g + 1, g + 1 + R.nx))
        p_out = p[sl]
    finally:
        R.close()
    np.savez_compressed(OUT / f"con2prim_{name}.npz", u_in=u_in, u_out=u_out, p_out=p_out, cells=cells)


def dust(nr):
    cfg = config_from_par(DUST.format(nr=nr))
    R = oracle.Reference(cfg)
    try:
        R.set_time(0.0, cfg.final_time)
        t, dt, n = R.step(10 ** 8)
        u, p = R.get_cons(), R.get_prim()
        ng = R.ng
    finally:
        R.close()
    x_min, dx = cfg.r_min, (cfg.r_max - cfg.r_min) / nr
    r = x_min + (np.arange(1, nr + 1) - 0.5) * dx
    exact = dust_exact_density(r, cfg.bh_mass, cfg.dust_accretion_constant)
    norms = dust_norms(p, exact, nr, cfg.ny, ng)
    np.savez_compressed(OUT / f"dust_{nr}.npz", u=u, p_row=p[ng + 1, ng + 1], t=t, n_steps=n, norms=norms)
    print(f"dust nr={nr}: {n} steps, L1/L2/Linf = {norms}")


if __name__ == "__main__":
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    which = sys.argv[1:] or STEP_CASES + ["sod_final", "michel_c2"]
    for w in which:
        if w == "sod_final":
            sod_final()
        elif w == "michel_c2":
            michel_c2(256)
        elif w.startswith("michel_c2_"):
            michel_c2(int(w.split("_")[-1]))
        elif w == "con2prim":
            for c in CON2PRIM_CASES:
                con2prim_fixture(c)
        elif w.startswith("dust_"):
            dust(int(w.split("_")[-1]))
        else:
            step_case(w)
        print("wrote", w, flush=True)
