"""Batch of independent problems (grhd_batch_*) on the GPU.

A problem's state after n steps in a batch must equal -- bit for bit -- its
state after n steps alone (grhd_advance): the batched kernels run the same
device code on each problem's own blocks, every launch form and face tiling
the autotuner may pick computes the same arithmetic, and each problem keeps
its own dt and time (grhd2.f90:179-184).  Tolerance: exact equality.
"""
import numpy as np
import pytest

from grhd_amd.batch import Batch, CapacityError, footprint
from grhd_amd.params import config_from_par
from grhd_amd.solver import Solver
from tests.cases import CASES, NOISY
from tests.helpers import seeded_noise

pytestmark = pytest.mark.gpu

MIX = ["sod_c1", "fm3d_ks_mp5_hllc_sensor", "michel_ef_log", "khi_weno3_sensor", "jet_cylindrical",
       "fm_sagittal_ef_log", "shu_tvd_minmod_sensor", "offaxis_minkowski_spherical_log"]
NSTEPS = 7


def _cfg(name, precision=None):
    text = CASES[name] + (f"precision = {precision}\n" if precision else "")
    return config_from_par(text)


def _initial(S, name):
    S.initial_conditions()
    if name in NOISY:
        kind, amp, seed = NOISY[name]
        p = seeded_noise(S.get_primitives(), S.nx, S.ny, S.nz, S.ng, kind, amp, seed)
        S.set_primitives(p)


def _alone(name, nsteps, precision=None):
    """State after nsteps steps of grhd_advance (fewer if the problem's final
    time comes first; a preset final_time of 0 takes none)."""
    with Solver(_cfg(name, precision)) as S:
        _initial(S, name)
        n, _ = S.advance(max_steps=nsteps)
        return S.get_conserved(), S.get_primitives(), S.t, n


def _batched(names, nsteps, variant=0, precision=None, chunks=1, pacing=1):
    out = {}
    with Batch(capacity=len(names) + 2) as B:
        B.set_variant(variant)
        B.set_pacing(pacing)
        slots = {}
        for n in names:
            slots[n] = B.add(_cfg(n, precision))
            _initial(B.problem(slots[n]), n)
        per = nsteps // chunks
        for c in range(chunks):
            B.step(per if c < chunks - 1 else nsteps - per * (chunks - 1))
        for n in names:
            view = B.problem(slots[n])
            done, t, steps = B.status(slots[n])
            out[n] = (view.get_conserved(), view.get_primitives(), t, steps)
        info = B.info()
    return out, info


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
def test_every_launch_form_matches_alone(variant):
    print(f"batch variant {variant}: {MIX}", flush=True)
    res, info = _batched(MIX, NSTEPS, variant=variant, chunks=2)
    assert info["variant"] == variant
    for n in MIX:
        u, p, t, na = _alone(n, NSTEPS)
        ub, pb, tb, steps = res[n]
        assert steps == na and tb == t, (n, steps, na, tb, t)
        assert np.array_equal(ub, u), n
        assert np.array_equal(pb, p), n


def test_autotuned_batch_matches_alone():
    print("batch autotuned", flush=True)
    res, info = _batched(MIX, 60, variant=0)  # the tuner times 5 forms x 2 x 2 steps first
    assert 1 <= info["variant"] <= 5
    for n in MIX:
        u, _, t, _ = _alone(n, 60)
        assert np.array_equal(res[n][0], u), n
        assert res[n][2] == t


def test_mixed_precision_problems_match_alone():
    names = ["fm3d_ks_mp5_hllc_sensor", "sod_c1", "michel_ef_log"]
    print("batch mixed precision", flush=True)
    res, _ = _batched(names, NSTEPS, variant=1, precision="mixed")
    for n in names:
        u, _, _, _ = _alone(n, NSTEPS, precision="mixed")
        assert np.array_equal(res[n][0], u), n


@pytest.mark.parametrize("variant", [1, 3])
def test_paced_batch_matches_alone(variant):
    """With pacing, groups whose steps are cheaper take more steps per chunk
    (up to 8 per step of the most expensive group): beside a 512^2
    Kelvin-Helmholtz problem (a step of about 190 us by the batch's estimate)
    the Michel and shock-tube groups (about 54 us) take three steps per chunk
    step.  Each problem's state equals its state alone after the steps it
    took, bit for bit."""
    big = CASES["khi_weno3_sensor"].replace("nx = 32", "nx = 512").replace("nz = 32", "nz = 512")
    texts = {"khi_512": big, "michel_ef_log": CASES["michel_ef_log"], "sod_c1": CASES["sod_c1"]}
    print(f"paced batch variant {variant}", flush=True)
    res = {}
    with Batch(capacity=4) as B:
        B.set_variant(variant)
        B.set_pacing(8)
        slots = {n: B.add(config_from_par(t)) for n, t in texts.items()}
        B.step(4)
        B.step(3)
        for n in texts:
            res[n] = (B.problem(slots[n]).get_conserved(), B.status(slots[n])[2])
    taken = {n: res[n][1] for n in texts}
    print("  steps taken", taken, flush=True)
    assert taken["khi_512"] == NSTEPS
    assert taken["michel_ef_log"] == 3 * NSTEPS and taken["sod_c1"] == 3 * NSTEPS
    for n, text in texts.items():
        with Solver(config_from_par(text)) as S:
            S.initial_conditions()
            na, _ = S.advance(max_steps=taken[n])
            assert na == taken[n]
            assert np.array_equal(res[n][0], S.get_conserved()), n


def test_pacing_needs_a_threefold_gap():
    """Beside a 400^2 Kelvin-Helmholtz problem (about 134 us by the estimate,
    2.5 times the small groups' 54 us) no group is paced: near the latency
    floor the estimate cannot separate such steps.  Every problem takes the
    chunk's steps and matches its run alone."""
    mid = CASES["khi_weno3_sensor"].replace("nx = 32", "nx = 400").replace("nz = 32", "nz = 400")
    texts = {"khi_400": mid, "michel_ef_log": CASES["michel_ef_log"], "sod_c1": CASES["sod_c1"]}
    print("pacing threshold", flush=True)
    res = {}
    with Batch(capacity=4) as B:
        B.set_variant(1)
        B.set_pacing(8)
        slots = {n: B.add(config_from_par(t)) for n, t in texts.items()}
        B.step(NSTEPS)
        for n in texts:
            res[n] = (B.problem(slots[n]).get_conserved(), B.status(slots[n])[2])
    assert {n: res[n][1] for n in texts} == {n: NSTEPS for n in texts}
    for n, text in texts.items():
        with Solver(config_from_par(text)) as S:
            S.initial_conditions()
            S.advance(max_steps=NSTEPS)
            assert np.array_equal(res[n][0], S.get_conserved()), n


def test_problems_stop_at_their_final_time():
    """A short problem finishes inside a chunk while a long one continues;
    the short one matches its own full run, final-step clamp included."""
    short = config_from_par(CASES["sod_c1"].replace("nx = 400", "nx = 100").replace("final_time = 0.4",
                                                                                      "final_time = 0.05"))
    long = _cfg("michel_ef_log")
    with Solver(short) as S:
        S.initial_conditions()
        n_alone, _ = S.advance()
        u_alone = S.get_conserved()
    with Batch(capacity=4) as B:
        a = B.add(short)
        b = B.add(long)
        B.step(n_alone + 10)
        done, t, steps = B.status(a)
        assert done and steps == n_alone and t == 0.05
        assert np.array_equal(B.problem(a).get_conserved(), u_alone)
        assert B.status(b)[2] == n_alone + 10 and not B.status(b)[0]


def test_admission_limits():
    cfg = _cfg("sod_c1")
    need = footprint(cfg)
    with Batch(capacity=2, hbm_budget=3 * need) as B:
        B.add(cfg)
        B.add(cfg)
        with pytest.raises(CapacityError):
            B.add(cfg)          # no slot
    with Batch(capacity=8, hbm_budget=need + need // 2) as B:
        s = B.add(cfg)
        with pytest.raises(CapacityError):
            B.add(cfg)          # budget
        B.remove(s)
        B.add(cfg)              # freed memory is reusable
