"""Device diagnostics: the mass monitor (grhd_total_mass, grhd2.f90:283-301)
and the convergence norms (grhd_convergence_norms, output.f90:1144-1258).

The mass is a device reduction whose summation order differs from the
reference's serial loop, so it is held to 1e-13 relative against an exactly
rounded host sum of the same per-cell terms u(D) dx dy dz.  The norms are
computed in the reference's order from the device's density row; they are
pinned against the reference build's own norms of the Michel C2 golden state
(tests/golden/michel_c2_256.npz), against the host restatement of the dust
and Michel branches, and by the identity that the advected wave's initial data
is its exact solution at t = 0.
"""
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from grhd_amd._lib import GrhdError
from grhd_amd.params import config_from_par
from grhd_amd.solver import Solver
from tests.cases import CASES, DUST, MICHEL_C2
from tests.helpers import dust_exact_density, dust_norms, michel_norms

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("name", ["sod_c1", "michel_ef_log", "fm3d_ks_mp5_hllc_sensor", "khi_weno3_sensor"])
def test_total_mass_matches_host_sum(name):
    cfg = config_from_par(CASES[name])
    print("mass", name, flush=True)
    O = oracle.Oracle(cfg)
    dx, dy, dz = O.grid()[:3]
    O.close()
    with Solver(cfg) as S:
        S.initial_conditions()
        S.advance(cfg.final_time, 5)
        u = S.get_conserved()
        m = S.total_mass()
    host = math.fsum((u[..., 0] * dx * dy * dz).ravel())
    assert m > 0.0
    assert abs(m - host) <= 1e-13 * abs(host), (m, host)


def test_michel_norms_match_reference_golden():
    """The reference build's calc_convergence_norms of its final state at
    t = 500 (nr = 256): the same state uploaded to the device gives the same
    norms (the exact solution differs from the reference's by the 1e-13 of the
    initial data, which moves the 1e-5 norms by about 1e-9 relative)."""
    g = np.load(GOLDEN / "michel_c2_256.npz")
    cfg = config_from_par(MICHEL_C2.format(nr=256))
    with Solver(cfg) as S:
        S.initial_conditions()
        p = S.get_primitives()
        ng = S.ng
        p[ng + 1, ng + 1] = g["p_row"]
        S.put_primitives(p)
        norms, im = S.convergence_norms(float(g["t"]))
    assert np.allclose(norms, g["norms"], rtol=1e-7, atol=0), (norms, g["norms"])
    assert int(0.10 * 256) <= im <= int(0.95 * 256)


def test_michel_norms_match_host_restatement():
    cfg = config_from_par(MICHEL_C2.format(nr=256))
    with Solver(cfg) as S:
        S.initial_conditions()
        exact = S.get_primitives()
        S.advance(500.0, 200)
        p = S.get_primitives()
        norms, _ = S.convergence_norms(0.0)
        ng = S.ng
    ref = michel_norms(p, exact, cfg.nx, cfg.ny, ng)
    assert np.allclose(norms, ref, rtol=1e-12, atol=0), (norms, ref)
    assert norms[0] > 0.0


def test_dust_norms_match_host_restatement():
    cfg = config_from_par(DUST.format(nr=128))
    with Solver(cfg) as S:
        S.initial_conditions()
        S.advance(500.0, 100)
        p = S.get_primitives()
        norms, im = S.convergence_norms(0.0)
        ng = S.ng
    r = cfg.r_min + (np.arange(1, cfg.nx + 1) - 0.5) * ((cfg.r_max - cfg.r_min) / cfg.nx)
    ref = dust_norms(p, dust_exact_density(r, cfg.bh_mass, cfg.dust_accretion_constant), cfg.nx, 1, ng)
    assert np.allclose(norms, ref, rtol=1e-12, atol=1e-300), (norms, ref)
    assert norms[2] > 0.0 and int(0.10 * 128) <= im <= int(0.95 * 128)


def test_advected_wave_norms():
    """The initial data is the exact cell average at t = 0, so every norm is
    zero there; after n steps the error is the scheme's, small and non-zero,
    and L1 <= L2 <= Linf."""
    cfg = config_from_par(CASES["conv_weno5_periodic"])
    with Solver(cfg) as S:
        S.initial_conditions()
        n0, _ = S.convergence_norms(0.0)
        assert n0 == (0.0, 0.0, 0.0)
        dt = S.compute_dt()
        S.step(dt, 20)
        n1, _ = S.convergence_norms(20 * dt)
        wrong, _ = S.convergence_norms(0.0)
    assert 0.0 < n1[0] <= n1[1] <= n1[2] < 1e-2 * cfg.convergence_amplitude
    # the exact solution moves with the wave: measured at t = 0 the error is larger
    assert wrong[0] > 10.0 * n1[0]


def test_norms_of_other_problems_are_unsupported():
    with Solver(config_from_par(CASES["sod_c1"])) as S:
        S.initial_conditions()
        with pytest.raises(GrhdError) as e:
            S.convergence_norms(0.0)
    assert e.value.code == 6


@pytest.mark.parametrize("name", ["fm3d_ks_mp5_hllc_sensor", "fm3d_ef"])
def test_observers_match_reference(name, tmp_path):
    """The device's global diagnostics, azimuthal modes and accretion rate of
    the reference's own state after three steps equal the reference's
    calc_global_diagnostics, calc_ppi_modes and calc_m_dot (output.f90:951-1125):
    counts, minima and maxima exactly, sums to 1e-12 relative (the reduction
    order differs), the modes' phases to 1e-9 absolute, the accretion rate to
    the 8 digits the reference writes."""
    if not oracle.reference_available():
        pytest.skip("reference build absent")
    cfg = config_from_par(CASES[name])
    R = oracle.Reference(cfg)
    try:
        R.set_time(0.0, 1.0e9)
        R.step(3)
        p, u = R.get_prim(), R.get_cons()
        scheme = {4: "mp5", 5: "weno5"}[cfg.reconstruction] + ("_hllc" if cfg.riemann_solver == 2 else "_hlle")
        ref = R.observers(str(tmp_path), scheme)
    finally:
        R.close()
    print("observers", name, flush=True)
    with Solver(cfg) as S:
        S.initial_conditions()
        S.put_primitives(p)
        S.put_conserved(u)
        g = S.global_diagnostics()
        modes = S.ppi_modes()
        mdot, probe = S.accretion_rate()
    gr = ref["global_diagnostics"]
    sums = [g["mass"], g["disk_mass"], g["angular_momentum"]]
    assert np.allclose(sums, gr[1:4], rtol=1e-12, atol=1e-300), (sums, gr[1:4])
    assert [g["rho_min"], g["p_min"], g["rho_max"], g["p_max"]] == list(gr[4:8])
    assert g["v2_max"] == pytest.approx(gr[8], rel=1e-14)
    assert (g["atmosphere_cells"], g["invalid_cells"]) == (int(gr[9]), int(gr[10]))
    mr = ref["ppi_modes"]
    assert np.allclose(modes[:1], mr[1:2], rtol=1e-12)
    for m in range(4):
        re, im, amp, ph = modes[1 + 4 * m: 5 + 4 * m]
        rre, rim, ramp, rph = mr[2 + 4 * m: 6 + 4 * m]
        scale = mr[1]
        assert abs(re - rre) <= 1e-12 * scale and abs(im - rim) <= 1e-12 * scale
        assert abs(amp - ramp) <= 1e-11
        assert abs(ph - rph) <= 1e-9 or amp < 1e-12
    # m_dot.dat holds 8 significant digits (E15.7, output.f90:989)
    assert mdot == pytest.approx(ref["m_dot"][1], rel=1e-6)
    assert probe >= 1
