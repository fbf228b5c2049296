"""The C oracle against the compiled reference (oracle/_ref), bit for bit.

Both are built with FP contraction off and use the same libm, so every stage
of the update -- boundaries, CFL dt, prim_to_cons, calc_rhs, full SSP-RK3 steps
with con2prim -- must agree exactly on every case.  This pins the oracle
restatement of evolution.f90, reconstruction.f90, fluxes.f90, equations.f90,
metrics.f90 and conditions.f90 (boundary part) to the reference itself.
"""
import numpy as np
import pytest

import oracle
from grhd_amd.params import config_from_par
from tests.cases import CASES, NOISY
from tests.helpers import seeded_noise

pytestmark = pytest.mark.skipif(not oracle.reference_available(),
                                reason="oracle/_ref not built (make -C oracle ref)")


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference_bitwise(name):
    cfg = config_from_par(CASES[name])
    R = oracle.Reference(cfg)
    try:
        O = oracle.Oracle(cfg)
        p = R.get_prim()
        if cfg.problem == 7:
            O.cache_injector(p)
        if name in NOISY:
            kind, amp, seed = NOISY[name]
            p = seeded_noise(p, O.nx, O.ny, O.nz, O.ng, kind, amp, seed)
            R.set_prim(p)
        # initial conserved state (grhd2.f90:126-134)
        u = O.prim_to_cons(p)
        assert np.array_equal(u, R.get_cons())
        # boundaries
        po = p.copy()
        O.boundaries(po)
        R.boundaries()
        assert np.array_equal(po, R.get_prim())
        # dt and rhs
        assert O.dt() == R.dt()
        assert np.array_equal(O.rhs(po), R.rhs())
        # three full steps through the reference loop
        R.set_time(0.0, cfg.final_time if cfg.final_time > 0 else 1.0)
        t_r, dt_r, _ = R.step(3)
        t_o, dt_o, n = O.run(u, po, 3, t=0.0, t_final=cfg.final_time if cfg.final_time > 0 else 1.0)
        assert t_o == t_r and dt_o == dt_r
        assert np.array_equal(u, R.get_cons())
        assert np.array_equal(po, R.get_prim())
    finally:
        R.close()
