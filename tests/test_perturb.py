"""Seeded perturbation (grhd_perturb) against the injection rules of
inject_pressure_noise / inject_density_noise (conditions.f90:603-636,
730-768): only interior cells with rho > 0.05 rho_max change; the factor
f = 1 + A (2 U - 1) lies in [1 - A, 1 + A]; pressure noise scales p by f;
density noise scales rho by f and p by f^Gamma; velocities and ghost cells
are untouched; U is rebuilt from the perturbed primitives.  Type 3 (no
reference counterpart) adds A (2 U - 1) to v^z of the same cells, the random
part of the Kelvin-Helmholtz seed perturbation.  The deviates come
from a counter-based generator keyed by (seed, cell) -- the reference's
runtime random_number stream is not reproduced (parity unpinned for the
deviates themselves) -- so the same seed gives the same field and different
seeds give different fields.
"""
import numpy as np
import pytest

from grhd_amd.params import config_from_par
from grhd_amd.solver import Solver
from tests.cases import CASES

pytestmark = pytest.mark.gpu

A = 1.0e-2


def _perturbed(kind, seed, name="fm3d_ks_mp5_hllc_sensor"):
    cfg = config_from_par(CASES[name])
    with Solver(cfg) as S:
        S.initial_conditions()
        p0 = S.get_primitives()
        S.perturb(kind, A, seed)
        p1 = S.get_primitives()
        u1 = S.get_conserved()
        S.set_primitives(p1)
        u_ref = S.get_conserved()
        ng, gam = S.ng, cfg.adiabatic_index
    return p0, p1, u1, u_ref, ng, gam


def _interior(a, ng):
    return a[ng + 1:-ng, ng + 1:-ng, ng + 1:-ng]


@pytest.mark.parametrize("kind", [1, 2])
def test_injection_rules(kind):
    print(f"perturb kind {kind}", flush=True)
    p0, p1, u1, u_ref, ng, gam = _perturbed(kind, 3435)
    i0, i1 = _interior(p0, ng), _interior(p1, ng)
    rho0 = i0[..., 0]
    hot = rho0 > 0.05 * rho0.max()
    assert hot.any() and (~hot).any()
    # untouched: cold cells, velocities, ghosts
    assert np.array_equal(i1[~hot], i0[~hot])
    assert np.array_equal(i1[..., 2:], i0[..., 2:])
    mask = np.ones(p0.shape[:3], bool)
    mask[ng + 1:-ng, ng + 1:-ng, ng + 1:-ng] = False
    assert np.array_equal(p1[mask], p0[mask])
    if kind == 1:
        assert np.array_equal(i1[..., 0], i0[..., 0])
        f = i1[..., 1][hot] / i0[..., 1][hot]
    else:
        f = i1[..., 0][hot] / i0[..., 0][hot]
        assert np.allclose(i1[..., 1][hot], i0[..., 1][hot] * f ** gam, rtol=1e-13, atol=0)
    assert np.all(f >= 1 - A - 1e-15) and np.all(f <= 1 + A + 1e-15)
    assert np.std(f) > 0.3 * A / np.sqrt(3)       # uniform on [1-A, 1+A]: std A/sqrt(3)
    assert abs(np.mean(f) - 1.0) < 5 * A / np.sqrt(3 * f.size)
    # U = P2C(p) of the perturbed primitives
    assert np.array_equal(u1, u_ref)


def test_seeded_and_reproducible():
    _, a, _, _, _, _ = _perturbed(2, 3435)
    _, b, _, _, _, _ = _perturbed(2, 3435)
    _, c, _, _, _, _ = _perturbed(2, 3436)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)


def test_transverse_velocity_noise():
    print("perturb kind 3 on khi_weno3_sensor", flush=True)
    p0, p1, u1, u_ref, ng, _ = _perturbed(3, 3435, "khi_weno3_sensor")
    i0, i1 = _interior(p0, ng), _interior(p1, ng)
    rho0 = i0[..., 0]
    hot = rho0 > 0.05 * rho0.max()
    assert hot.any()
    assert np.array_equal(i1[~hot], i0[~hot])
    assert np.array_equal(i1[..., :4], i0[..., :4])
    d = i1[..., 4][hot] - i0[..., 4][hot]
    assert np.all(np.abs(d) <= A * (1 + 1e-12))
    assert np.std(d) > 0.3 * A / np.sqrt(3)
    # ghosts of the active x and z axes untouched (the case has ny = 1: the
    # reference layout's y ghost slots hold copies of the single row)
    row0, row1 = p0[:, ng + 1], p1[:, ng + 1]
    mask = np.ones(row0.shape[:2], bool)
    mask[ng + 1:-ng, ng + 1:-ng] = False
    assert np.array_equal(row1[mask], row0[mask])
    assert np.array_equal(u1, u_ref)
