"""Host logic of the serving loop on the CPU, with a stand-in plan.

The stand-in advances each problem by a fixed dt per step and adds 1 to every
cell of its state per step, so a request's final state equals its own step
count only if its planes were carried over intact through every re-plan, and
its time state (t, steps) only if ah_plan_time_export/import kept it.
"""
import contextlib
import math

import numpy as np
import pytest
import torch

from astrea_amd import serve

DT = 0.05


class FakePlan:
    live = 0          # plans alive at once (class-wide)
    max_cells = 0     # most cells held by plans alive at once
    device_ic = 0     # initial states built by init_quadrants

    def __init__(self, dims, subgrid, solver, timestep, problems, **kw):
        self.problems = problems
        self.nv, self.varidx = 4, [0, 1, 2, 4]
        sizes = [p["n0"] * p["n1"] for p in problems]
        self.offsets = [int(x) for x in np.cumsum([0] + sizes[:-1])]
        self.total = sum(sizes)
        self.state = torch.zeros((self.nv, self.total), dtype=torch.float64)
        self.times = [(0.0, 0.0, 0, 1, 1, 0, 0) for _ in problems]
        FakePlan.live += self.total
        FakePlan.max_cells = max(FakePlan.max_cells, FakePlan.live)

    def time_export(self):
        return list(self.times)

    def time_import(self, times):
        self.times = [tuple(t) for t in times]

    def step(self, n):
        for i, p in enumerate(self.problems):
            t, dt, steps, idx, active, failed, par = self.times[i]
            o, cells = self.offsets[i], p["n0"] * p["n1"]
            for _ in range(n):
                if not active:
                    break
                t, steps = t + DT, steps + 1
                self.state[:, o:o + cells] += 1.0
                if t >= p["t_end"] - 1e-12:
                    active = 0
            self.times[i] = (t, DT, steps, idx, active, failed, par)

    def step_graph(self, n, k):
        self.step(n * k)

    def load(self, grids):
        for i, g in enumerate(grids):
            o, cells = self.offsets[i], self.problems[i]["n0"] * self.problems[i]["n1"]
            flat = torch.as_tensor(g).reshape(-1, 8)
            for v, c in enumerate(self.varidx):
                self.state[v, o:o + cells] = flat[:, c]

    def init_quadrants(self, i, states):
        o, cells = self.offsets[i], self.problems[i]["n0"] * self.problems[i]["n1"]
        FakePlan.device_ic += 1
        self.state[:, o:o + cells] = 0.0

    def reset(self):
        self.times = [(0.0, 0.0, 0, 1, 1, 0, 0) for _ in self.problems]

    def run(self, chunk):
        while any(t[4] for t in self.times):
            self.step(chunk)
        return self.status()

    def status(self):
        return [dict(t=t, dt=dt, steps=s, active=bool(a), failed=bool(f)) for t, dt, s, _, a, f, _ in self.times]

    def close(self):
        FakePlan.live -= self.total


class _Stream:
    def __init__(self, *a, **k):
        pass

    def synchronize(self):
        pass

    def wait_stream(self, other):
        pass


@pytest.fixture
def fake_gpu(monkeypatch):
    FakePlan.live = FakePlan.max_cells = FakePlan.device_ic = 0
    monkeypatch.setattr(serve, "Plan", FakePlan)
    monkeypatch.setattr(torch.cuda, "Stream", _Stream)
    monkeypatch.setattr(torch.cuda, "stream", lambda s: contextlib.nullcontext())
    monkeypatch.setattr(torch.cuda, "current_stream", lambda *a, **k: _Stream())


def _requests(spec):
    out = []
    for rid, (n, sub, t_end, arrival) in enumerate(spec):
        out.append(serve.Request(rid, n, sub, arrival, t_end, None, torch.zeros((n, n, 8), dtype=torch.float64)))
    return out


SPEC = [(16, "plm", .5, 0.0), (32, "ppm", .25, 0.0), (16, "ppm", 1.0, 0.0), (8, "plm", .1, 0.0),
        (32, "plm", .3, 0.0), (8, "ppm", .6, 0.0), (16, "plm", .2, 0.0)]


def test_server_finishes_every_request_with_its_own_state(fake_gpu):
    srv = serve.Server(hbm_budget=1 << 40, chunk=3, device="cpu", keep_results=True, bytes_per_cell=1)
    out = srv.run(_requests(SPEC), realtime=False)
    assert out["requests"] == len(SPEC) and out["failed"] == 0
    for r in out["done"]:
        steps = math.ceil(r.t_end / DT - 1e-9)
        assert r.status["steps"] == steps, (r.rid, r.status)
        # every cell of every plane advanced exactly once per step of this request
        assert torch.all(r.result == steps)
    assert out["rebuilds"] > 1


def test_admission_respects_the_hbm_budget(fake_gpu):
    # room for 1.5 of the 32^2 requests: at most one of them resident at a time
    budget = int(1.5 * 32 * 32)
    srv = serve.Server(hbm_budget=budget, chunk=2, device="cpu", keep_results=True, bytes_per_cell=1)
    out = srv.run(_requests(SPEC), realtime=False)
    assert out["requests"] == len(SPEC)
    assert FakePlan.max_cells <= budget
    for r in out["done"]:
        assert torch.all(r.result == r.status["steps"])


def test_served_equals_fifo(fake_gpu):
    a = serve.Server(hbm_budget=1 << 40, chunk=4, device="cpu", keep_results=True, bytes_per_cell=1).run(
        _requests(SPEC), realtime=False)
    b = serve.run_fifo(_requests(SPEC), chunk=5, realtime=False, keep_results=True, device="cpu")
    ra = {r.rid: r for r in a["done"]}
    for r in b["done"]:
        assert r.status["steps"] == ra[r.rid].status["steps"]
        assert torch.equal(r.result.reshape(-1), ra[r.rid].result.reshape(-1))


def test_request_larger_than_the_budget_is_refused(fake_gpu):
    srv = serve.Server(hbm_budget=100, chunk=2, device="cpu", bytes_per_cell=1)
    with pytest.raises(ValueError, match="does not fit"):
        srv.run(_requests([(16, "plm", .2, 0.0)]), realtime=False)


def test_active_cell_cap_bounds_the_batch(fake_gpu, monkeypatch):
    cap = 32 * 32 + 16 * 16
    seen = []
    advance = serve._Group.advance

    def watch(self, chunk, use_graph=True):
        seen.append(sum(r.n * r.n for g in srv.groups.values() for r in g.reqs))
        return advance(self, chunk, use_graph)

    monkeypatch.setattr(serve._Group, "advance", watch)
    srv = serve.Server(hbm_budget=1 << 40, chunk=2, device="cpu", keep_results=True, bytes_per_cell=1,
                       max_active_cells=cap)
    out = srv.run(_requests(SPEC), realtime=False)
    assert out["requests"] == len(SPEC)
    assert seen and max(seen) <= cap
    for r in out["done"]:
        assert torch.all(r.result == r.status["steps"])


def test_backfill_admits_small_requests_past_a_held_large_one(fake_gpu, monkeypatch):
    """With backfill, a small request queued behind a large one that the cell
    cap holds back is admitted at once; the large ones keep their arrival
    order, and the cap still bounds every batch."""
    spec = [(32, "plm", .5, 0.0), (32, "ppm", .5, 0.0), (8, "plm", .1, 0.0), (32, "plm", .3, 0.0),
            (8, "ppm", .1, 0.0)]
    cap = 32 * 32 + 2 * 8 * 8
    seen = []
    advance = serve._Group.advance

    def watch(self, chunk, use_graph=True):
        seen.append(sum(r.n * r.n for g in srv.groups.values() for r in g.reqs))
        return advance(self, chunk, use_graph)

    monkeypatch.setattr(serve._Group, "advance", watch)
    order = {}
    for backfill in (False, True):
        seen.clear()
        srv = serve.Server(hbm_budget=1 << 40, chunk=2, device="cpu", keep_results=True, bytes_per_cell=1,
                           max_active_cells=cap, backfill=backfill)
        out = srv.run(_requests(spec), realtime=False)
        assert out["requests"] == len(spec) and max(seen) <= cap
        for r in out["done"]:
            assert torch.all(r.result == r.status["steps"])
        order[backfill] = {r.rid: r.admitted_at for r in out["done"]}
    plain, bf = order[False], order[True]
    # without backfill the small requests wait for the large one ahead of them
    assert plain[2] > plain[1] and plain[4] > plain[1]
    # with it they start beside the first request, before the held one
    assert bf[2] < bf[1] and bf[4] < bf[1]
    assert bf[1] < bf[3]


def test_requests_without_grids_start_on_the_device(fake_gpu):
    reqs = _requests(SPEC)
    for r in reqs:
        r.grid, r.states = None, [(1., 0., 0., 1.)] * 4
    srv = serve.Server(hbm_budget=1 << 40, chunk=3, device="cpu", keep_results=True, bytes_per_cell=1)
    out = srv.run(reqs, realtime=False)
    assert out["requests"] == len(SPEC) and FakePlan.device_ic == len(SPEC)
    for r in out["done"]:
        assert torch.all(r.result == r.status["steps"])
