"""Analytic-solution errors of the device path against the reference's.

Dust (problem 8: cold free fall onto Schwarzschild in EF coordinates, r in
[1, 51], WENO5 + HLLE + sensor, t = 500) at nr = 128, 256, 512: from the
initial data the device reaches the reference's final state and the same
L1/L2/Linf density errors against the stationary analytic solution
(output.f90:1214-1235), stored in tests/golden/dust_<nr>.npz.

C2 Michel accretion (EF, log r, nr = 256 and 1024, WENO5 + HLLE, t = 500,
1e-3 seeded density noise): the device run from the reference's perturbed
initial state reaches the reference's final state and the same L1/L2/Linf
density errors against the analytic solution (output.f90:1189-1213), stored
in tests/golden/michel_c2_<nr>.npz by tests/golden/make_golden.py.  nr = 4096
(945k steps) has no reference value (the reference needs about 19 CPU-hours for
it): there the test holds the step count of the clamped constant-dt loop and
the convergence of the error below the reference's at nr = 1024.

The analytic solution is the problem's initial data; the device's initial
data equals the reference's to 1e-13 (test_gpu_parity), so the norms are
taken against grhd_initial_conditions' primitives.
"""
from pathlib import Path

import numpy as np
import pytest

from grhd_amd.params import config_from_par
from tests.cases import DUST, MICHEL_C2
from tests.helpers import dust_exact_density, dust_norms, michel_norms, rel_errors

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_michel_golden_norms_consistent():
    """The stored norms follow from the stored final row and the reference's
    analytic initial data (conditions.f90:284-410)."""
    import oracle
    if not oracle.reference_available():
        pytest.skip("reference build absent")
    g = np.load(GOLDEN / "michel_c2_256.npz")
    cfg = config_from_par(MICHEL_C2.format(nr=256))
    R = oracle.Reference(cfg)
    try:
        exact = R.get_prim()
        ng = R.ng
    finally:
        R.close()
    p = exact.copy()
    p[ng + 1, ng + 1] = g["p_row"]
    assert np.array_equal(michel_norms(p, exact, cfg.nx, cfg.ny, ng), g["norms"])
    # The discretisation error dominates the noise residue at t = 500.
    assert 1e-6 < g["norms"][0] < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nr", [256, 1024])
def test_michel_c2(nr, precision):
    from grhd_amd.solver import Solver
    g = np.load(GOLDEN / f"michel_c2_{nr}.npz")
    cfg = config_from_par(MICHEL_C2.format(nr=nr) + f"precision = {precision}\n")
    print(f"michel C2 nr={nr} {precision}: {int(g['n_steps'])} steps", flush=True)
    with Solver(cfg) as S:
        S.initial_conditions()
        exact = S.get_primitives()
        S.set_primitives(g["p0"])
        n, _ = S.advance()
        t = S.t
        u = S.get_conserved()
        p = S.get_primitives()
        ng = S.ng
    norms = michel_norms(p, exact, cfg.nx, cfg.ny, ng)
    print(f"  device norms {norms}  reference {g['norms']}", flush=True)
    if precision == "fp64":
        assert n == int(g["n_steps"]) and t == float(g["t"])
        l1, li = rel_errors(u, g["u"])
        assert np.all(l1 <= 1e-9) and np.all(li <= 1e-7), (l1, li)
        assert np.allclose(norms, g["norms"], rtol=1e-4, atol=0)
    else:
        # FP32 faces: the error against the analytic solution stays that of
        # the scheme (within 10 % of the FP64 reference's L1) or, where the
        # scheme's error is below it, the FP32 floor of a smooth stationary
        # flow (a few 1e-6 in rho).
        assert abs(norms[0] - g["norms"][0]) <= max(0.1 * g["norms"][0], 5.0e-6), (norms, g["norms"])


def _clamped_steps(dt, t_final):
    """Steps of the reference time loop at a constant dt (grhd2.f90:179-232):
    t advances by dt, and the step that would pass t_final is clamped to it."""
    t, n = 0.0, 0
    while t < t_final:
        step = dt if t + dt <= t_final else t_final - t
        t = t + step
        n += 1
    return n, t


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_michel_c2_4096(precision):
    """C2 at nr = 4096 (944,742 steps, about 80 s in FP64).  The reference's
    value here is unpinned (it needs about 19 CPU-hours), so this holds what
    the mathematics fixes: the step count of the constant-dt loop with the
    final clamp, an FP64 L1 density error below the reference's at nr = 1024
    (the scheme converges: 1.36e-6 there), and a mixed-precision error within
    the FP32 floor of a smooth stationary flow (1e-6)."""
    from grhd_amd.solver import Solver
    from tests.helpers import seeded_noise
    g1024 = np.load(GOLDEN / "michel_c2_1024.npz")
    cfg = config_from_par(MICHEL_C2.format(nr=4096) + f"precision = {precision}\n")
    with Solver(cfg) as S:
        S.initial_conditions()
        exact = S.get_primitives()
        S.set_primitives(seeded_noise(exact, S.nx, S.ny, S.nz, S.ng, 2, 1.0e-3, 3435))
        dt = S.compute_dt()
        n_expected, t_expected = _clamped_steps(dt, cfg.final_time)
        print(f"michel C2 nr=4096 {precision}: {n_expected} steps", flush=True)
        n, _ = S.advance()
        t = S.t
        p = S.get_primitives()
        ng = S.ng
    norms = michel_norms(p, exact, cfg.nx, cfg.ny, ng)
    print(f"  device norms {norms}  reference at nr=1024 {g1024['norms']}", flush=True)
    assert n == n_expected and t == t_expected == cfg.final_time
    assert np.all(np.isfinite(norms))
    if precision == "fp64":
        assert norms[0] < 0.5 * float(g1024["norms"][0]), norms
    else:
        assert norms[0] <= 1.0e-6, norms


DUST_NR = [128, 256, 512]


@pytest.mark.parametrize("nr", DUST_NR)
def test_dust_golden_norms_follow_from_final_row(nr):
    g = np.load(GOLDEN / f"dust_{nr}.npz")
    cfg = config_from_par(DUST.format(nr=nr))
    ng = 3
    r = cfg.r_min + (np.arange(1, nr + 1) - 0.5) * ((cfg.r_max - cfg.r_min) / nr)
    exact = dust_exact_density(r, cfg.bh_mass, cfg.dust_accretion_constant)
    p = np.zeros((1, 1 + 2 * ng + 1, nr + 2 * ng + 1, 5))
    p[0, ng + 1] = g["p_row"]
    p = np.concatenate([np.zeros_like(p)] * (ng + 1) + [p], axis=0)
    assert np.array_equal(dust_norms(p, exact, nr, 1, ng), g["norms"])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nr", DUST_NR)
def test_dust(nr, precision):
    from grhd_amd.solver import Solver
    g = np.load(GOLDEN / f"dust_{nr}.npz")
    cfg = config_from_par(DUST.format(nr=nr) + f"precision = {precision}\n")
    print(f"dust nr={nr} {precision}: {int(g['n_steps'])} steps", flush=True)
    with Solver(cfg) as S:
        S.initial_conditions()
        n, _ = S.advance()
        t = S.t
        u = S.get_conserved()
        p = S.get_primitives()
        ng = S.ng
    r = cfg.r_min + (np.arange(1, nr + 1) - 0.5) * ((cfg.r_max - cfg.r_min) / nr)
    norms = dust_norms(p, dust_exact_density(r, cfg.bh_mass, cfg.dust_accretion_constant), nr, 1, ng)
    print(f"  device norms {norms}  reference {g['norms']}", flush=True)
    assert n == int(g["n_steps"]) and t == float(g["t"])
    if precision == "fp64":
        l1, li = rel_errors(u, g["u"])
        assert np.all(l1 <= 1e-9) and np.all(li <= 1e-7), (l1, li)
        assert np.allclose(norms, g["norms"], rtol=1e-3, atol=0)
    else:
        # FP32 faces: the flux difference F(i+1) - F(i) of a smooth stationary
        # flow cancels most digits of two FP32 fluxes, which leaves an error
        # floor of a few 1e-6 in rho: at coarse grids the L1 error stays that
        # of the scheme, at fine grids it is the floor (nr 256: 4.8e-6 against
        # the FP64 path's 2.2e-7)
        assert norms[0] <= max(2.0 * g["norms"][0], 1.0e-5), (norms, g["norms"])
