"""The .par surface: presets and parser against the reference's own parser.

Values resolved by grhd_amd.params must equal what the reference reads from the
same file (grid extents, ghost depth, method ids, spacings, CFL, floors).
"""
import math

import pytest

import oracle
from grhd_amd import params
from tests.cases import CASES


@pytest.mark.skipif(not oracle.reference_available(), reason="reference build not present")
@pytest.mark.parametrize("name", sorted(CASES))
def test_resolved_config_matches_reference(name):
    cfg = params.config_from_par(CASES[name])
    R = oracle.Reference(par_text=CASES[name])
    try:
        d = R.dims
        s = R.scalars()
    finally:
        R.close()
    assert (d[0], d[1], d[2], d[3]) == (cfg.nx, cfg.ny, cfg.nz, cfg.nghost)
    assert d[4] == cfg.problem and d[5] == cfg.reconstruction and d[6] == cfg.riemann_solver
    assert d[7] == int(cfg.shock_sensor)
    O = oracle.Oracle(cfg)
    g = O.grid()
    assert list(g[:9]) == list(s[:9])
    assert s[9] == cfg.cfl and s[10] == cfg.final_time and s[11] == cfg.adiabatic_index
    assert s[12] == cfg.bh_mass and s[13] == cfg.spin
    assert s[14] == 1e-10 and s[15] == 1e-13


@pytest.mark.parametrize("text,msg", [
    ("nx = 4\n", "missing mandatory parameter 'problem'"),
    ("problem = sod\nnx = 4\nnx = 5\n", "duplicated"),
    ("problem = sod\nfoo = 1\n", "unknown parameter"),
    ("problem = sod\nfm_inner_radius = 6\n", "does not belong"),
    ("problem = sod\nreconstruction = weno7\n", "unknown reconstruction"),
    ("problem = michel\nspin = 0.5\n", "Eddington-Finkelstein needs spin = 0"),
    ("problem = fishbone_3d\nnz = 7\n", "fishbone_3d"),
    ("problem = sod\nnx = 4 = 5\n", "more than one"),
])
def test_rejections(text, msg):
    with pytest.raises(params.ParameterError, match=msg):
        params.config_from_par(text)


def test_pi_fractions_and_case_insensitivity():
    cfg = params.config_from_par("PROBLEM = Fishbone_3D\ntheta_min = 0\nTheta_Max = 2/2\n"
                                 "phi_max = 2  # comment\nRECONSTRUCTION = MP5\n")
    assert cfg.y_max == math.pi and cfg.z_max == 2 * math.pi and cfg.reconstruction == 4
    assert cfg.nghost == 3
