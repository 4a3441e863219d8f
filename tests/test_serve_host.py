"""The serving loop's scheduling (heatflow_mpm.serve.Server) on the CPU, with a fake backend and a
simulated clock: every request runs exactly its substeps, admission follows arrival order and
never exceeds the device budget or the slot count, requests that arrive while others run join
the running batch, and max_active=1 is the plain one-at-a-time form."""
import numpy as np
import pytest

from heatflow_mpm import GridConfig
from heatflow_mpm.serve import Request, Server, footprint_estimate, summarize


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, s):
        self.t += s


class FakeSim:
    def __init__(self, req):
        self.req, self.steps, self.n = req, 0, req.kwargs["grid"].n


class FakeBackend:
    """Each batched substep costs 0.1 ms plus 1 ms per million cells of the batch."""

    def __init__(self, clock):
        self.clock, self.batches, self.max_seen = clock, [], 0

    def prepare(self, req):
        return None

    def create(self, req, prepared):
        return FakeSim(req)

    def n_of(self, sim):
        return sim.n

    def bytes(self, sim):
        return footprint_estimate(sim.n, len(sim.req.kwargs["particles"]))

    def make_batch(self, sims):
        self.batches.append([s.req.rid for s in sims])
        self.max_seen = max(self.max_seen, len(sims))
        return list(sims)

    def set_batch(self, batch, sims):
        batch[:] = list(sims)
        if sims:
            self.batches.append([s.req.rid for s in sims])
            self.max_seen = max(self.max_seen, len(sims))

    def substep(self, batch, dt):
        for s in batch:
            s.steps += 1
        self.clock.t += 1e-4 + 1e-3 * sum(s.n ** 3 for s in batch) / 1e6

    def collect(self, sim, keep):
        return {"steps": sim.steps}, {}

    def close_batch(self, batch):
        pass

    def close_sim(self, sim):
        pass


def _reqs(k=12, seed=1):
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    for i in range(k):
        t += float(rng.exponential(0.01))
        n = int(rng.choice([16, 24, 32]))
        out.append(Request(i, dict(particles=np.zeros((int(rng.integers(0, 100)), 3), np.float32),
                                   grid=GridConfig(L=n * 0.01, h=0.01)), int(rng.integers(3, 40)), t))
    return out


def _serve(reqs, **kw):
    clock = Clock()
    be = FakeBackend(clock)
    srv = Server(backend=be, clock=clock, sleep=clock.sleep, keep_results=True, **kw)
    res = srv.run(reqs)
    return srv, be, res


def test_every_request_runs_exactly_its_substeps():
    reqs = _reqs()
    srv, be, res = _serve(reqs)
    assert [r.rid for r in res] == [q.rid for q in reqs]
    for r, q in zip(res, reqs):
        assert r.fields["steps"] == q.substeps and r.finished >= r.admitted >= r.arrival - 1e-12


def test_budget_and_slots_are_respected_and_admission_is_in_arrival_order():
    reqs = _reqs(20)
    budget = 3 * footprint_estimate(32, 100)
    srv, be, res = _serve(reqs, hbm_budget=budget, max_active=4)
    assert max(e[4] for e in srv.log) <= budget
    assert be.max_seen <= 4
    admits = [e[2] for e in srv.log if e[1] == "admit"]
    assert admits == sorted(admits)


def test_late_arrivals_join_a_running_batch():
    reqs = [Request(0, dict(particles=np.zeros((0, 3), np.float32), grid=GridConfig(L=0.32, h=0.01)), 50, 0.0),
            Request(1, dict(particles=np.zeros((0, 3), np.float32), grid=GridConfig(L=0.16, h=0.01)), 5, 0.005)]
    srv, be, res = _serve(reqs)
    assert [0, 1] in be.batches                      # request 1 joined while request 0 ran
    assert res[1].finished < res[0].finished         # and left before it


def test_one_slot_is_the_plain_one_at_a_time_form():
    reqs = _reqs(8)
    srv, be, res = _serve(reqs, max_active=1)
    assert be.max_seen == 1
    fin = [r.finished for r in res]
    assert fin == sorted(fin)
    s = summarize(res)
    assert s["requests"] == 8 and s["p99_s"] >= s["p50_s"] > 0


def test_a_request_larger_than_the_budget_is_refused():
    reqs = _reqs(3)
    with pytest.raises(ValueError, match="does not fit"):
        _serve(reqs, hbm_budget=1000)


def test_multigrid_footprint_covers_the_coarse_levels():
    """Six fp32 arrays over the coarse levels (n/2, n/4, ... down to at most 8 per side)."""
    for n in (16, 24, 48, 96, 128, 192):
        coarse, m = 0, n
        while m > 8 and m % 2 == 0:
            m //= 2
            coarse += m ** 3
        assert footprint_estimate(n, 1000, cg="mg") - footprint_estimate(n, 1000) >= 24 * coarse


def test_lanes_keep_size_classes_apart_and_serve_every_request():
    """Size-class lanes: every request runs its substeps, a lane's batches hold its size class only,
    and the shared budget is never exceeded."""
    clock = Clock()
    be = FakeBackend(clock)
    reqs = _reqs(16, seed=3)
    budget = 6 * footprint_estimate(32, 100)
    srv = Server(be, hbm_budget=budget, keep_results=True, clock=clock, sleep=clock.sleep, lanes=[(24, 8), (32, 1)])
    res = srv.run(reqs)
    assert [r.rid for r in res] == [q.rid for q in reqs]
    for r, q in zip(res, reqs):
        assert r.fields["steps"] == q.substeps
    by_rid = {q.rid: q.kwargs["grid"].n for q in reqs}
    for b in be.batches:
        sizes = {by_rid[rid] for rid in b}
        assert sizes <= {16, 24} or (sizes == {32} and len(b) == 1), b
    assert max(e[4] for e in srv.log) <= budget
    with pytest.raises(ValueError, match="larger than every lane"):
        Server(be, lanes=[(16, 4)], clock=clock, sleep=clock.sleep).run(reqs)
