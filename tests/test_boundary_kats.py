"""Known-answer tests of the ghost-zone fills, taken from the reference's own unit tests.

* tests/test_polar_boundaries.f90:13-111 -- fishbone_3d: the trans-polar fill
  (ghost row 1-g at phi index k holds row g at k + nz/2, v^theta negated) and
  periodic phi; fishbone_sagittal: reflection across the poles with v^theta
  negated.  The interior is the test's pattern 1e5*c + 1e3*i + 10*j + k and the
  ghosts start at -huge; every check is exact equality.
* tests/test_case_corrections.f90:266-294 -- the shift-aware radial outflow of
  the Fishbone-Moncrief families: an inner ghost v^x is capped at beta^x/alpha
  of cell 1 and an outer ghost floored at beta^x/alpha of cell nx, so that no
  fluid enters through either radial face (conditions.f90:916-947, 996-1019).
  The reference test sets alpha = 0.8, beta = 0.2 by hand; here the lapse and
  shift are the analytic Kerr-Schild ones at the cell centres.

Each check runs on the oracle (CPU) and, marked gpu, through the C ABI.
"""
import math

import numpy as np
import pytest

import oracle
from grhd_amd.params import config_from_par

HUGE = np.finfo(np.float64).max
DE, PR, VX, VY, VZ = range(5)   # eq_de, eq_pr, eq_vx, eq_vy, eq_vz (variables.f90:26-30)

POLAR_3D = """problem = fishbone_3d
nx = 3
ny = 4
nz = 6
reconstruction = weno3
"""
POLAR_SAGITTAL = """problem = fishbone_sagittal
nx = 3
ny = 4
reconstruction = weno3
"""
SHIFT_AWARE = """problem = fishbone_equatorial
metric = kerr_schild
logarithmic_r = true
nx = 8
nz = 4
r_min = 1.8
r_max = 20
reconstruction = weno3
"""


def _pattern(pshape, nx, ny, nz, ng):
    """test_polar_boundaries.f90:47-61, in the (k, j, i, component) layout."""
    p = np.full(pshape, -HUGE)
    k, j, i, c = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1),
                             np.arange(1, 6), indexing="ij")
    p[ng + 1:ng + 1 + nz, ng + 1:ng + 1 + ny, ng + 1:ng + 1 + nx, :] = 1.0e5 * c + 1.0e3 * i + 10.0 * j + k
    return p


def _at(p, ng, i, j, k):
    """p(:, i, j, k); the arrays span -ng..n+ng on every axis."""
    return p[k + ng, j + ng, i + ng]


def _check_3d_poles(p, nx, ny, nz, ng):
    """test_polar_boundaries.f90:64-91."""
    assert ng == 2 and nz % 2 == 0
    for k in range(1, nz + 1):
        ks = (k - 1 + nz // 2) % nz + 1
        for g in range(1, ng + 1):
            north = _at(p, ng, 1, g, ks).copy()
            north[VY] = -north[VY]
            south = _at(p, ng, 1, ny + 1 - g, ks).copy()
            south[VY] = -south[VY]
            assert np.array_equal(_at(p, ng, 1, 1 - g, k), north), ("north", k, g)
            assert np.array_equal(_at(p, ng, 1, ny + g, k), south), ("south", k, g)
    for g in range(1, ng + 1):
        assert np.array_equal(_at(p, ng, 1, 2, 1 - g), _at(p, ng, 1, 2, nz + 1 - g))
        assert np.array_equal(_at(p, ng, 1, 2, nz + g), _at(p, ng, 1, 2, g))


def _check_sagittal_poles(p, ny, ng):
    """test_polar_boundaries.f90:94-111."""
    for g in range(1, ng + 1):
        north = _at(p, ng, 1, g, 1).copy()
        north[VY] = -north[VY]
        south = _at(p, ng, 1, ny + 1 - g, 1).copy()
        south[VY] = -south[VY]
        assert np.array_equal(_at(p, ng, 1, 1 - g, 1), north), ("north", g)
        assert np.array_equal(_at(p, ng, 1, ny + g, 1), south), ("south", g)


def _shift_aware_input(O):
    """Unit density and pressure, v = 0, except v^x at cell 1 above the cap
    and v^x at cell nx below the floor (test_case_corrections.f90:276-284).
    Returns the state and the two expected ghost velocities beta^x/alpha."""
    dx, x_min = O.grid()[0], O.grid()[3]
    th = math.pi / 2
    m1, _, _ = O.metric_point(x_min + 0.5 * dx, th)
    mn, _, _ = O.metric_point(x_min + (O.nx - 0.5) * dx, th)
    cap, floor = m1[1] / m1[0], mn[1] / mn[0]
    assert cap > 0.0 and floor > 0.0   # Kerr-Schild: inward shift at both radii
    p = np.zeros(O.pshape)
    ng = O.ng
    p[..., DE] = 1.0
    p[..., PR] = 1.0
    p[ng + 1:ng + 1 + O.nz, ng + 1, ng + 1, VX] = 1.5 * cap
    p[ng + 1:ng + 1 + O.nz, ng + 1, ng + O.nx, VX] = 0.0
    return p, cap, floor


def _check_shift_aware(p, O, cap, floor, rtol):
    ng, nx, nz = O.ng, O.nx, O.nz
    for k in range(1, nz + 1):
        for g in range(1, ng + 1):
            inner = _at(p, ng, 1 - g, 1, k)
            outer = _at(p, ng, nx + g, 1, k)
            assert abs(inner[VX] - cap) <= rtol * cap, (k, g, inner[VX], cap)
            assert abs(outer[VX] - floor) <= rtol * floor, (k, g, outer[VX], floor)
            assert inner[DE] == 1.0 and inner[PR] == 1.0 and outer[DE] == 1.0


# ------------------------------------------------------------------ oracle
def test_oracle_trans_polar_3d():
    O = oracle.Oracle(config_from_par(POLAR_3D))
    p = _pattern(O.pshape, O.nx, O.ny, O.nz, O.ng)
    O.boundaries(p)
    _check_3d_poles(p, O.nx, O.ny, O.nz, O.ng)


def test_oracle_sagittal_poles():
    O = oracle.Oracle(config_from_par(POLAR_SAGITTAL))
    assert O.nz == 1
    p = _pattern(O.pshape, O.nx, O.ny, O.nz, O.ng)
    O.boundaries(p)
    _check_sagittal_poles(p, O.ny, O.ng)


def test_oracle_shift_aware_outflow():
    O = oracle.Oracle(config_from_par(SHIFT_AWARE))
    p, cap, floor = _shift_aware_input(O)
    O.boundaries(p)
    _check_shift_aware(p, O, cap, floor, rtol=0.0)


def test_shift_aware_check_catches_a_plain_copy():
    """A plain outflow copy (the shock-tube rule) leaves v^x = 1.5 cap in the
    inner ghosts and 0 in the outer ones: the check above rejects it."""
    O = oracle.Oracle(config_from_par(SHIFT_AWARE))
    p, cap, floor = _shift_aware_input(O)
    ng = O.ng
    for g in range(1, ng + 1):
        p[:, :, ng + 1 - g] = p[:, :, ng + 1]
        p[:, :, ng + O.nx + g] = p[:, :, ng + O.nx]
    with pytest.raises(AssertionError):
        _check_shift_aware(p, O, cap, floor, rtol=1e-3)


# ------------------------------------------------------------------ device
@pytest.mark.gpu
@pytest.mark.parametrize("par,check", [(POLAR_3D, "3d"), (POLAR_SAGITTAL, "sagittal")])
def test_device_polar_fills(par, check):
    from grhd_amd.solver import Solver
    with Solver(config_from_par(par)) as S:
        p = _pattern(S.pshape, S.nx, S.ny, S.nz, S.ng)
        S.put_primitives(p)
        S.boundaries()
        q = S.get_primitives()
    if check == "3d":
        _check_3d_poles(q, S.nx, S.ny, S.nz, S.ng)
    else:
        _check_sagittal_poles(q, S.ny, S.ng)


@pytest.mark.gpu
def test_device_shift_aware_outflow():
    from grhd_amd.solver import Solver
    cfg = config_from_par(SHIFT_AWARE)
    O = oracle.Oracle(cfg)
    p, cap, floor = _shift_aware_input(O)
    with Solver(cfg) as S:
        S.put_primitives(p)
        S.boundaries()
        q = S.get_primitives()
    # the device evaluates alpha and beta with its own sqrt and division
    _check_shift_aware(q, O, cap, floor, rtol=1e-14)


def test_polar_check_catches_an_unshifted_or_unsigned_fill():
    """A pole fill without the k + nz/2 shift, or without the v^theta sign
    flip, fails the 3D check."""
    O = oracle.Oracle(config_from_par(POLAR_3D))
    ng, ny, nz = O.ng, O.ny, O.nz
    for shift, sign in ((0, -1.0), (nz // 2, 1.0)):
        p = _pattern(O.pshape, O.nx, ny, nz, ng)
        O.boundaries(p)
        for k in range(1, nz + 1):
            ks = (k - 1 + shift) % nz + 1
            for g in range(1, ng + 1):
                p[k + ng, 1 - g + ng] = p[ks + ng, g + ng]
                p[k + ng, 1 - g + ng, :, VY] *= sign
        with pytest.raises(AssertionError):
            _check_3d_poles(p, O.nx, ny, nz, ng)
