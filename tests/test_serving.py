"""The continuous-batching request loop (grhd_amd/serving.py).

CPU: the scheduling logic against a stand-in batch whose problems finish
after a fixed number of steps -- admission in arrival order, first-fit under
the slot and HBM limits, requests arriving while others run join the batch,
every request completes once, latencies are completion - arrival.
GPU: a served mix of problems returns for each request the state it has
when run alone (bit-identical), and the loop reports p50/p99 beside
throughput.
"""
import numpy as np
import pytest

from grhd_amd import params as P
from grhd_amd.serving import Request, Server, request_mix


class FakeView:
    def __init__(self, cells):
        self.n_cells = cells

    def get_conserved(self):
        return np.zeros(1)

    def perturb(self, *a):
        pass


class FakeBatch:
    """Problem r finishes after cfg.nx steps; records admissions and steps."""

    def __init__(self, capacity):
        self.capacity = capacity
        self.slots = {}
        self.log = []
        self.next = 0
        self.clock = 0.0

    def add(self, cfg):
        s = self.next
        self.next += 1
        self.slots[s] = [cfg, 0]
        self.log.append(("add", s, self.clock, len(self.slots)))
        return s

    def step(self, n):
        for v in self.slots.values():
            v[1] = min(v[1] + n, v[0].nx)
        self.clock += 1.0  # one time unit per chunk
        self.log.append(("step", len(self.slots)))

    def status(self, s):
        cfg, st = self.slots[s]
        return st >= cfg.nx, 0.0, st

    def problem(self, s):
        return FakeView(self.slots[s][0].nx)

    def remove(self, s):
        del self.slots[s]


def _req(rid, nx, arrival, size=1):
    cfg = P.preset(1).replace(nx=nx, ny=1, nz=size)
    return Request(rid, cfg, arrival, f"r{rid}")


def _server(batch, budget, chunk=10):
    return Server(batch, footprint=lambda cfg: cfg.nz, hbm_budget=budget, chunk=chunk,
                  clock=lambda: batch.clock, sleep=lambda s: setattr(batch, "clock", batch.clock + s))


def test_all_requests_complete_with_latencies():
    B = FakeBatch(capacity=4)
    reqs = [_req(i, nx=10 * (1 + i % 3), arrival=0.5 * i) for i in range(10)]
    rep = _server(B, budget=100).run(reqs)
    assert sorted(c.rid for c in rep.completed) == list(range(10))
    for c in rep.completed:
        assert c.finished >= c.admitted >= c.arrival
        assert c.steps == reqs[c.rid].cfg.nx
    s = rep.summary()
    assert s["requests"] == 10 and s["p99_s"] >= s["p50_s"] > 0


def test_slot_limit_and_continuous_admission():
    B = FakeBatch(capacity=2)
    reqs = [_req(0, 100, 0.0), _req(1, 10, 0.0), _req(2, 10, 0.0), _req(3, 10, 0.0)]
    rep = _server(B, budget=100).run(reqs)
    assert max(e[3] for e in B.log if e[0] == "add") <= 2
    # request 2 joins while request 0 (long) is still running
    done = {c.rid: c for c in rep.completed}
    assert done[2].admitted > 0.0 and done[2].admitted < done[0].finished


def test_budget_first_fit():
    B = FakeBatch(capacity=8)
    # footprints (nz): 6, 6, 1 with budget 8: the third is admitted beside the
    # first while the second waits
    reqs = [_req(0, 50, 0.0, size=6), _req(1, 10, 0.0, size=6), _req(2, 10, 0.0, size=1)]
    rep = _server(B, budget=8).run(reqs)
    done = {c.rid: c for c in rep.completed}
    assert done[0].admitted == done[2].admitted == 0.0
    assert done[1].admitted >= done[0].finished - 1e-12


def test_request_larger_than_budget_is_rejected():
    with pytest.raises(ValueError):
        _server(FakeBatch(2), budget=3).run([_req(0, 10, 0.0, size=5)])


def test_request_mix_is_seeded_and_valid():
    a = request_mix(20, rate=10.0, seed=7)
    b = request_mix(20, rate=10.0, seed=7)
    assert [r.label for r in a] == [r.label for r in b]
    assert all(r2.arrival > r1.arrival for r1, r2 in zip(a, a[1:]))
    kinds = {r.label.split()[0] for r in a}
    assert {"michel", "kelvin_helmholtz"} & kinds
    # Kelvin-Helmholtz requests carry seeded v^z noise (type 3), distinct per request
    khi = [r for r in a if r.label.startswith("kelvin_helmholtz")]
    assert khi and all(r.perturb is not None and r.perturb[:2] == (3, 1.0e-2) for r in khi)
    assert len({r.perturb[2] for r in khi}) == len(khi)
    assert all(r.perturb is None for r in a if not r.label.startswith("kelvin_helmholtz"))


@pytest.mark.gpu
def test_served_results_equal_alone():
    from grhd_amd.batch import Batch, footprint
    from grhd_amd.solver import Solver
    reqs = request_mix(10, rate=200.0, seed=11, time_scale=0.02, shock_nx=(200, 400),
                       michel_nr=(256,), khi_n=(64,))
    for r in reqs:
        r.want_state = True
    print("serving:", [r.label for r in reqs], flush=True)
    with Batch(capacity=6) as B:
        rep = Server(B, footprint, hbm_budget=1 << 34, chunk=16).run(reqs)
    s = rep.summary()
    print("serving summary", s, flush=True)
    assert s["requests"] == len(reqs) and s["p99_s"] >= s["p50_s"] > 0
    assert rep.peak_active > 1
    for c in rep.completed:
        with Solver(reqs[c.rid].cfg) as S:
            S.initial_conditions()
            if reqs[c.rid].perturb is not None:
                S.perturb(*reqs[c.rid].perturb)
            n, _ = S.advance()
            assert n == c.steps, c.label
            assert np.array_equal(S.get_conserved(), c.state), c.label


@pytest.mark.gpu
def test_served_results_equal_alone_at_workload_sizes():
    """The served workload at its stated sizes in one mix: shock tubes at
    nx = 3200 and 6400, Michel accretion at nr = 4096 on Kerr-Schild with
    a = 0.97 and at nr = 1024 with a = 0, Kelvin-Helmholtz at 256^2 and
    1024^2 (Gamma = 4/3, v = +-0.5, seeded v^z noise), arriving while others
    run.  Final times are cut to 1 % of the presets' (64 steps of the nx = 6400
    tube, 10,656 of the nr = 4096 Michel run, 246 of the 1024^2
    Kelvin-Helmholtz run) so that it runs in well under a minute; every result
    equals the same problem run alone through grhd_advance, bit for bit."""
    from grhd_amd.batch import Batch, footprint
    from grhd_amd.serving import Request, _kelvin_helmholtz, _michel, _shock_tube
    from grhd_amd.solver import Solver
    scale = 0.01
    specs = [("shu_osher nx=6400", _shock_tube(3, 6400), None),
             ("michel nr=4096 a=0.97", _michel(4096, 0.97), None),
             ("kelvin_helmholtz 1024x1024", _kelvin_helmholtz(1024), (3, 1.0e-2, 17)),
             ("strong_shock nx=3200", _shock_tube(2, 3200), None),
             ("michel nr=1024 a=0", _michel(1024, 0.0), None),
             ("kelvin_helmholtz 256x256", _kelvin_helmholtz(256), (3, 1.0e-2, 18))]
    reqs = [Request(i, cfg.replace(final_time=cfg.final_time * scale), 0.05 * i, label, perturb=pert,
                    want_state=True) for i, (label, cfg, pert) in enumerate(specs)]
    print("serving at workload sizes:", [r.label for r in reqs], flush=True)
    with Batch(capacity=8) as B:
        rep = Server(B, footprint, hbm_budget=1 << 34, chunk=32).run(reqs)
    s = rep.summary()
    print("serving summary", s, flush=True)
    assert s["requests"] == len(reqs) and rep.peak_active > 1
    for c in rep.completed:
        r = reqs[c.rid]
        with Solver(r.cfg) as S:
            S.initial_conditions()
            if r.perturb is not None:
                S.perturb(*r.perturb)
            n, _ = S.advance()
            assert n == c.steps, c.label
            assert np.array_equal(S.get_conserved(), c.state), c.label
        print(f"served {c.label}: {c.steps} steps, equal to the run alone", flush=True)
