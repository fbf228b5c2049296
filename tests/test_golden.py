"""Golden vectors written by the reference build (tests/golden/make_golden.py).

The oracle must reproduce them bit for bit (same rounding sequence, same
libm); the HIP path within the FP64 tolerances of test_gpu_parity.py.  These
tests need neither the reference sources nor its build.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle
from grhd_amd.params import config_from_par
from tests.cases import CASES
from tests.helpers import read_set_mask, rel_errors

GOLD = Path(__file__).resolve().parent / "golden"
STEP_CASES = sorted(p.stem for p in GOLD.glob("*.npz") if p.stem in CASES)


def _oracle_for(name, p0):
    cfg = config_from_par(CASES[name])
    O = oracle.Oracle(cfg)
    if cfg.problem == 7:
        O.cache_injector(p0)
    return cfg, O


@pytest.mark.parametrize("name", STEP_CASES)
def test_oracle_reproduces_golden_steps(name):
    g = np.load(GOLD / f"{name}.npz")
    cfg, O = _oracle_for(name, g["p0"])
    p = g["p0"].copy()
    u = O.prim_to_cons(p)
    assert O.dt() == float(g["dt"])
    t_final = cfg.final_time if cfg.final_time > 0 else 1.0
    O.run(u, p, 1, t=0.0, t_final=t_final)
    assert np.array_equal(u, g["u1"]) and np.array_equal(p, g["p1"])
    t, _, _ = O.run(u, p, 9, t=float(O.dt()), t_final=t_final)
    assert np.array_equal(u, g["u10"])


def test_oracle_reproduces_sod_c1_final_state():
    g = np.load(GOLD / "sod_c1_final.npz")
    cfg, O = _oracle_for("sod_c1", g["p0"])
    p = g["p0"].copy()
    u = O.prim_to_cons(p)
    t, dt, n = O.run(u, p, 10 ** 6)
    assert n == int(g["n_steps"]) == 400 and t == float(g["t"])
    assert np.array_equal(u, g["u"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", STEP_CASES)
def test_gpu_reproduces_golden_steps(name):
    from grhd_amd.solver import Solver
    g = np.load(GOLD / f"{name}.npz")
    cfg = config_from_par(CASES[name])
    with Solver(cfg) as S:
        S.set_primitives(g["p0"])
        S.step(float(g["dt"]), 1)
        l1, li = rel_errors(S.get_conserved(), g["u1"])
        assert l1.max() <= 1e-12 and li.max() <= 1e-12, f"{l1} {li}"
        m = read_set_mask(S.nx, S.ny, S.nz, S.ng)
        l1, li = rel_errors(S.get_primitives()[m], g["p1"][m])
        assert l1.max() <= 1e-11 and li.max() <= 1e-9, f"{l1} {li}"


@pytest.mark.gpu
def test_gpu_sod_c1_final_state():
    from grhd_amd.solver import Solver
    g = np.load(GOLD / "sod_c1_final.npz")
    cfg = config_from_par(CASES["sod_c1"])
    with Solver(cfg) as S:
        S.set_primitives(g["p0"])
        n, _ = S.advance(cfg.final_time)
        assert n == 400 and S.t == float(g["t"])
        l1, li = rel_errors(S.get_conserved(), g["u"])
    assert l1.max() <= 1e-10, f"{l1} {li}"
