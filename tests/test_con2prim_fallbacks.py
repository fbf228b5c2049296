"""The con2prim fallbacks and the CFL error path, pinned to the reference.

cons_to_prim (equations.f90:94-205) falls back to Brent's method
(root_brent, equations.f90:343-424) only when rtsafe (root_safe_newton,
equations.f90:249-340) has not converged after 100 iterations
(equations.f90:337-339).  Ordinary flows never get there: rtsafe bisects on
[1, max(1e3, 2 W_guess)] with W_guess = (|S| + D) / D, so only states with
tau/D and |S|/D beyond about 1e24 need more than 100 halvings.  The fixtures
tests/golden/con2prim_<case>.npz (written by the reference build,
tests/golden/make_golden.py) plant such states, and NaN, infinite, negative-D
and negative-tau states that must reset to the atmosphere
(equations.f90:116-124, 186-204), among the initial conserved state of three
cases (Kerr-Schild 3D torus, Michel on Kerr-Schild a = 0.9, Minkowski Sod).

update_adaptive_dt (evolution.f90:198-275) error-stops on invalid metric data;
the C ABI returns GRHD_ERR_INVALID_METRIC instead.  A Minkowski spherical grid
out to r = 1e300 passes the reference's validation, and r^2 overflows there,
so gamma^thetatheta = 0 in every cell.
"""
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from grhd_amd.params import config_from_par
from tests.cases import CASES

ROOT = Path(__file__).resolve().parents[1]
GOLD = ROOT / "tests" / "golden"
C2P_CASES = ["fm3d_ks_mp5_hllc_sensor", "michel_ks_a090", "sod_mp5_hllc_sensor"]

HUGE_R = """problem = offaxis_blast
metric = minkowski
geometry = spherical
bh_mass = 0
nx = 4
ny = 4
nz = 4
r_min = 1
r_max = 1e300
offaxis_radial_center = 2
"""


# the reference's constant floors (initialization.f90:454-455)
RHO_FLOOR = 1.0e-10
P_FLOOR = RHO_FLOOR * 1.0e-3


def _atmosphere_cells(p_out):
    """Interior cells whose primitives are exactly the atmosphere state."""
    q = p_out.reshape(-1, 5)
    return int(np.sum((q[:, 0] == RHO_FLOOR) & (q[:, 1] == P_FLOOR) & (q[:, 2] == 0.0) &
                      (q[:, 3] == 0.0) & (q[:, 4] == 0.0)))


def _oracle_con2prim(name, u_in):
    cfg = config_from_par(CASES[name])
    O = oracle.Oracle(cfg)
    g = O.ng
    u = u_in.copy()
    p = np.zeros(O.pshape)
    oracle.brent_stats(reset=True)
    O.cons_to_prim(u, p)
    calls, iters = oracle.brent_stats(reset=True)
    sl = (slice(g + 1, g + 1 + O.nz), slice(g + 1, g + 1 + O.ny), slice(g + 1, g + 1 + O.nx))
    O.close()
    return cfg, u, p[sl], calls, iters


@pytest.mark.parametrize("name", C2P_CASES)
def test_oracle_reproduces_con2prim_fallback_fixture(name):
    g = np.load(GOLD / f"con2prim_{name}.npz")
    cfg, u, p, calls, iters = _oracle_con2prim(name, g["u_in"])
    assert np.array_equal(u, g["u_out"], equal_nan=True)
    assert np.array_equal(p, g["p_out"], equal_nan=True)
    # the planted hot states reach Brent and run its loop
    assert calls >= 20 and iters >= 20 * calls // 2, (calls, iters)
    # some of them come back from Brent with a root (finite, not the atmosphere) ...
    planted = g["u_in"].reshape(-1, 5)[g["cells"]]
    q = g["p_out"].reshape(-1, 5)[g["cells"]]
    hot = np.isfinite(planted).all(axis=1) & (planted[:, 1] > 1e20)
    ok = hot & (q[:, 0] != RHO_FLOOR)
    assert ok.sum() >= 10 and np.isfinite(q[ok]).all()
    # ... and the NaN, infinite and negative states end in the atmosphere
    bad = ~np.isfinite(planted).all(axis=1) | (planted[:, 0] < 0) | (planted[:, 1] < 0)
    assert bad.sum() >= 7
    assert np.all(q[bad, 0] == RHO_FLOOR) and np.all(q[bad, 1] == P_FLOOR)
    assert np.all(q[bad, 2:] == 0.0)
    ub = g["u_out"].reshape(-1, 5)[g["cells"]][bad]
    assert np.all(ub[:, 2:] == 0.0) and np.all(ub[:, 0] > 0.0) and np.all(ub[:, 1] > 0.0)


@pytest.mark.skipif(not oracle.reference_available(), reason="oracle/_ref not built")
def test_reference_cell_con2prim_matches_oracle_on_brent_states():
    """Cell by cell through the reference's own cons_to_prim entry
    (ref_cons_to_prim_cell), at two cells of the Kerr-Schild torus."""
    sys.path.insert(0, str(ROOT / "tests" / "golden"))
    from make_golden import con2prim_states  # noqa: E402
    cfg = config_from_par(CASES["fm3d_ks_mp5_hllc_sensor"])
    R = oracle.Reference(cfg)
    O = oracle.Oracle(cfg)
    try:
        oracle.brent_stats(reset=True)
        for (i, j, k) in [(2, 3, 1), (17, 8, 5)]:
            for s in con2prim_states():
                cr, qr = R.cons_to_prim_cell(s, i, j, k)
                co, qo = O.cons_to_prim_cell(s, i, j, k)
                assert np.array_equal(cr, co, equal_nan=True) and np.array_equal(qr, qo, equal_nan=True), s
        assert oracle.brent_stats(reset=True)[0] > 50
    finally:
        R.close()
        O.close()


def test_oracle_rejects_invalid_cfl_metric():
    O = oracle.Oracle(config_from_par(HUGE_R))
    assert O.dt() == -1.0
    O.close()


@pytest.mark.skipif(not oracle.reference_available(), reason="oracle/_ref not built")
def test_reference_error_stops_on_invalid_cfl_metric(tmp_path):
    """The reference ends with error stop 1 in update_adaptive_dt
    (evolution.f90:257-260) on the same grid (in a child process)."""
    par = tmp_path / "huge.par"
    par.write_text(HUGE_R)
    code = ("import sys; sys.path.insert(0, %r); import oracle; "
            "from grhd_amd.params import config_from_par; "
            "R = oracle.Reference(config_from_par(open(%r).read())); R.dt()") % (str(ROOT), str(par))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
    assert r.returncode == 1
    assert "invalid metric data in CFL cells" in r.stdout + r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name", C2P_CASES)
def test_gpu_con2prim_brent_and_atmosphere(name):
    """grhd_con2prim on the planted states: bit for bit the reference's
    output, the Brent counter equal to the oracle's Brent calls and the
    atmosphere counter equal to the cells left in the atmosphere state."""
    from grhd_amd.solver import Solver
    g = np.load(GOLD / f"con2prim_{name}.npz")
    cfg, _, _, calls, _ = _oracle_con2prim(name, g["u_in"])
    print(f"con2prim fallbacks {name}: {calls} Brent calls expected", flush=True)
    with Solver(cfg) as S:
        S.initial_conditions()
        c0 = S.counters()
        S.put_conserved(np.ascontiguousarray(g["u_in"]))
        S.con2prim()
        c1 = S.counters()
        u = S.get_conserved()
        p = S.get_primitives()
        ng = S.ng
    sl = (slice(ng + 1, ng + 1 + cfg.nz), slice(ng + 1, ng + 1 + cfg.ny), slice(ng + 1, ng + 1 + cfg.nx))
    assert np.array_equal(u, g["u_out"], equal_nan=True)
    assert np.array_equal(p[sl], g["p_out"], equal_nan=True)
    assert c1["brent_fallbacks"] - c0["brent_fallbacks"] == calls > 0
    assert c1["atmosphere_resets"] - c0["atmosphere_resets"] == _atmosphere_cells(g["p_out"])


@pytest.mark.gpu
def test_gpu_compute_dt_invalid_metric():
    from grhd_amd._lib import GRHD_ERR_INVALID_METRIC, GrhdError
    from grhd_amd.solver import Solver
    cfg = config_from_par(HUGE_R)
    with Solver(cfg) as S:
        with pytest.raises(GrhdError) as e:
            S.compute_dt()
    assert e.value.code == GRHD_ERR_INVALID_METRIC
    assert "invalid metric data in CFL cells" in str(e.value)
