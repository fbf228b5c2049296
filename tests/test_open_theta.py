"""The 3D torus with theta away from the poles (the measured 256 x 128 x 128
workload runs theta in [0.05 pi, 0.95 pi]).

The reference admits fishbone_3d only with theta = [0, pi]
(parameters.f90:577-583), where its trans-polar ghost fill
(conditions.f90:1032-1050) continues the grid across the pole.  A theta
boundary that is not a pole gets an outflow fill here instead (no reference
counterpart): a copy of the edge cell, rho and p floored, v^theta limited to
outflow (<= 0 at theta_min, >= 0 at theta_max; beta^theta = 0 for these
metrics, so this is the shift-aware limit of the reference's radial outflow).

- CPU: the oracle's fill against that rule, value by value, on a random state.
- GPU: the device fill against the oracle's, bit for bit; one RK3 step
  against the oracle at 1e-12; and, at the full 256 x 128 x 128 size, that an
  axisymmetric torus stays axisymmetric bit for bit (every phi plane equal)
  and finite over steps of the outflow boundary.
"""
import numpy as np
import pytest

import oracle
from grhd_amd import params
from tests.helpers import read_set_mask, rel_errors, seeded_noise

TORUS = """problem = fishbone_3d
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
"""
RHO_FLOOR = 1.0e-10
P_FLOOR = RHO_FLOOR * 1.0e-3


def open_theta_config(nx, ny, nz, tmin=0.05, tmax=0.95):
    text = TORUS.format(nx=nx, ny=ny, nz=nz, tmin=tmin, tmax=tmax)
    cfg = params.preset(12)
    for _, k, v in params.parse_par_lines(text):
        if k != "problem":
            params._apply_key(cfg, k, v, 12)
    cfg.extra["_allow_open_theta"] = True
    params.validate(cfg)
    return cfg


def test_open_theta_validation():
    cfg = open_theta_config(16, 8, 8)
    assert abs(cfg.y_min - 0.05 * np.pi) < 1e-15 and abs(cfg.y_max - 0.95 * np.pi) < 1e-15
    with pytest.raises(params.ParameterError):
        c = params.preset(12)
        c.y_min = 0.05 * np.pi
        params.validate(c)


@pytest.mark.parametrize("tmin,tmax", [(0.05, 0.95), (0.0, 0.9), (0.1, 1.0)])
def test_oracle_open_theta_fill_rule(tmin, tmax):
    cfg = open_theta_config(12, 8, 6, tmin, tmax)
    O = oracle.Oracle(cfg)
    ng, nx, ny, nz = O.ng, O.nx, O.ny, O.nz
    rng = np.random.default_rng(7)
    p = rng.uniform(0.5, 2.0, O.pshape)
    p[..., 2:] = rng.uniform(-0.3, 0.3, p[..., 2:].shape)
    # some edge cells below the floors, to see them floored in the ghosts
    p[ng + 1:ng + 1 + nz, ng + 1, :, 0] = 1e-14
    p[ng + 1:ng + 1 + nz, ng + ny, :, 1] = 1e-16
    q = p.copy()
    O.boundaries(q)
    O.close()
    I = slice(ng + 1, ng + 1 + nx)
    for k in range(ng + 1, ng + 1 + nz):
        for side, open_ in (("lo", tmin > 0), ("hi", tmax < 1)):
            edge = ng + 1 if side == "lo" else ng + ny
            mirror_k = (k - ng - 1 + nz // 2) % nz + ng + 1
            for g in range(1, ng + 1):
                jg = ng + 1 - g if side == "lo" else ng + ny + g
                got = q[k, jg, I]
                if open_:
                    e = p[k, edge, I]
                    assert np.array_equal(got[:, 0], np.maximum(e[:, 0], RHO_FLOOR))
                    assert np.array_equal(got[:, 1], np.maximum(e[:, 1], P_FLOOR))
                    lim = np.minimum(e[:, 3], 0.0) if side == "lo" else np.maximum(e[:, 3], 0.0)
                    assert np.array_equal(got[:, 3], lim)
                    assert np.array_equal(got[:, 2], e[:, 2]) and np.array_equal(got[:, 4], e[:, 4])
                else:  # a pole: trans-polar fill from phi + pi, v^theta reversed
                    js = ng + g if side == "lo" else ng + ny + 1 - g
                    src = p[mirror_k, js, I]
                    assert np.array_equal(got[:, [0, 1, 2, 4]], src[:, [0, 1, 2, 4]])
                    assert np.array_equal(got[:, 3], -src[:, 3])


@pytest.mark.gpu
def test_gpu_open_theta_fill_and_step_match_oracle():
    from grhd_amd.solver import Solver
    cfg = open_theta_config(32, 16, 16)
    print("open theta 32x16x16: fill and one RK3 step against the oracle", flush=True)
    with Solver(cfg) as S:
        S.initial_conditions()
        p = seeded_noise(S.get_primitives(), S.nx, S.ny, S.nz, S.ng, 2, 1.0e-2, 3435)
        S.set_primitives(p)
        O = oracle.Oracle(cfg)
        S.boundaries()
        p_o = p.copy()
        O.boundaries(p_o)
        m = read_set_mask(S.nx, S.ny, S.nz, S.ng)
        assert np.array_equal(S.get_primitives()[m], p_o[m])
        u = O.prim_to_cons(p)
        dt = O.dt()
        assert abs(S.compute_dt() - dt) <= 1e-14 * dt
        S.step(dt, 1)
        O.step(dt, u, p.copy())
        l1, li = rel_errors(S.get_conserved(), u)
        O.close()
    assert l1.max() <= 1e-12 and li.max() <= 1e-12, f"{l1} {li}"


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_gpu_open_theta_full_size_axisymmetry(precision):
    """256 x 128 x 128, theta in [0.05 pi, 0.95 pi], unperturbed torus: after
    four RK3 steps every phi plane is bit for bit the first one, and the state
    is finite (FP64 and mixed)."""
    from grhd_amd.solver import Solver, fishbone_unit_density_k
    cfg = open_theta_config(256, 128, 128).replace(precision=precision)
    cfg.fm_polytropic_constant = fishbone_unit_density_k(cfg)
    print(f"open theta 256x128x128 precision {precision}: 4 steps", flush=True)
    with Solver(cfg) as S:
        S.initial_conditions()
        S.step(S.compute_dt(), 4)
        u = S.get_conserved()
        p = S.get_primitives()
        ng = S.ng
    assert np.isfinite(u).all()
    assert np.all(u == u[:1])
    sl = (slice(ng + 1, ng + 1 + cfg.nz), slice(ng + 1, ng + 1 + cfg.ny), slice(ng + 1, ng + 1 + cfg.nx))
    q = p[sl]
    assert np.isfinite(q).all() and np.all(q[..., 0] >= RHO_FLOOR)
