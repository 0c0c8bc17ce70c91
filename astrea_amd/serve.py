"""Continuous-batching solve server for streams of 2D four-quadrant requests.

Requests (2D problems of different sizes, schemes and end times) arrive over
time.  Admitted requests keep their state resident on the device; requests
of one scheme share one batched plan, each with its own t and dt on the
device.  Admission is against an HBM budget: a request is admitted when the
re-plans it causes stay within it -- the groups re-plan one at a time, each
new plan built while its old one is alive, and the peak of that sequence
(computed in Server._replan_peak) must fit.  Whenever the active set of a scheme
changes (an admission or a completion) its plan is rebuilt and the running
requests' state planes and time states move over (device-to-device copies),
so new work joins the batch between chunks of steps instead of waiting for
the batch to drain.  Each scheme group runs on its own stream, so groups
overlap on the GPU.

``run_fifo`` is the plain comparison: one request at a time, FP64, eager
launches, in arrival order.
"""
import collections
import time
from dataclasses import dataclass, field

import torch

from . import ic
from .plan import Plan


@dataclass
class Request:
    rid: int
    n: int
    subgrid: str
    arrival: float
    t_end: float = .25
    states: list = None                # four-quadrant primitive states (rho, u, v, p)
    grid: torch.Tensor = None          # or an explicit initial conservative state (n, n, 8)
    admitted_at: float = None
    done_at: float = None
    status: dict = None
    result: torch.Tensor = None        # final packed state planes (nv, n*n)
    extra: dict = field(default_factory=dict)

    @property
    def latency(self):
        return None if self.done_at is None else self.done_at - self.arrival


def make_requests(nreq, seed, rate, sizes=(128, 256, 512, 1024), device="cuda:0", host_grids=False):
    """Seeded four-quadrant requests with Poisson arrivals at ``rate`` per
    second.  A request carries its four states only: its initial grid is
    built on the device when it is admitted (ah_plan_init_ic).  With
    ``host_grids`` each request also carries the host-pipeline grid."""
    import numpy as np
    specs = ic.quadrant_requests(nreq, seed, sizes)
    rng = np.random.default_rng(seed + 1)
    t = 0.0
    out = []
    for i, sp in enumerate(specs):
        t += float(rng.exponential(1.0 / rate))
        g = None
        if host_grids:
            g = ic.to_device(ic.quadrants(sp["n"], sp["states"]), 1.4, "edge", sp["subgrid"] == "ppm", device)
        out.append(Request(i, sp["n"], sp["subgrid"], t, .25, sp["states"], g))
    return out


def _initial_state(plan, j, r):
    """Problem j of ``plan`` starts from request r's initial state."""
    if r.grid is None:
        plan.init_quadrants(j, r.states)
        return
    o, cells = plan.offsets[j], r.n * r.n
    flat = r.grid.reshape(-1, 8)
    for v, c in enumerate(plan.varidx):
        plan.state[v, o:o + cells] = flat[:, c]


def _problem(r, cfl):
    return dict(n0=r.n, n1=r.n, boundary="edge", dx=1 / r.n, gamma=1.4, cfl=cfl, t_end=r.t_end)


class _Group:
    def __init__(self, subgrid, precision, compat, cfl, device):
        self.subgrid, self.precision, self.compat, self.cfl = subgrid, precision, compat, cfl
        self.device = torch.device(device)
        self.stream = torch.cuda.Stream(device=self.device)
        self.reqs = []
        self.plan = None
        self.stale = False

    def _new_plan(self, reqs):
        return Plan(2, self.subgrid, "hllc", "ssprk(3,3)", [_problem(r, self.cfl) for r in reqs],
                    precision=self.precision, varmask=0x17, compat=self.compat, device=self.device)

    def rebuild(self):
        """New plan over self.reqs; running requests keep state and time."""
        # new requests' initial states were written on the caller's stream
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(self.stream):
            old, old_reqs = self.plan, getattr(self, "_planned", [])
            if not self.reqs:
                if old:
                    old.close()
                self.plan, self._planned, self.stale = None, [], False
                return
            new = self._new_plan(self.reqs)
            times = []
            old_times = old.time_export() if old else []
            old_index = {r.rid: i for i, r in enumerate(old_reqs)}
            for j, r in enumerate(self.reqs):
                o_new = new.offsets[j]
                cells = r.n * r.n
                if r.rid in old_index:
                    i = old_index[r.rid]
                    o_old = old.offsets[i]
                    new.state[:, o_new:o_new + cells].copy_(old.state[:, o_old:o_old + cells])
                    times.append(old_times[i])
                else:
                    _initial_state(new, j, r)
                    active = 1 if (0.0 <= r.t_end and r.t_end != 0.0) else 0
                    times.append((0.0, 0.0, 0, 1, active, 0, 0))
            new.time_import(times)
            if old:
                old.close()
            self.plan, self._planned, self.stale = new, list(self.reqs), False

    def advance(self, chunk, use_graph=True):
        if self.plan is None:
            return
        with torch.cuda.stream(self.stream):
            if use_graph:
                self.plan.step_graph(chunk, 1)
            else:
                self.plan.step(chunk)

    def held(self, cost):
        """Bytes of the group's current plan: a re-plan builds the new plan
        while this one is alive (the running requests' planes are copied
        over), so admission has to leave room for both."""
        return sum(cost(r) for r in getattr(self, "_planned", []))

    def harvest(self, now, keep_results):
        """Finished requests (done or failed) leave the group."""
        if self.plan is None:
            return []
        with torch.cuda.stream(self.stream):
            st = self.plan.status()
        done = []
        for i, (r, s) in enumerate(zip(self._planned, st)):
            if not s["active"]:
                r.done_at = now()
                r.status = s
                if keep_results:
                    o = self.plan.offsets[i]
                    with torch.cuda.stream(self.stream):
                        r.result = self.plan.state[:, o:o + r.n * r.n].clone()
                    self.stream.synchronize()  # the caller reads it on its own stream
                done.append(r)
        if done:
            ids = {r.rid for r in done}
            self.reqs = [r for r in self.reqs if r.rid not in ids]
            self.stale = True
        return done


class Server:
    def __init__(self, hbm_budget=32 << 30, precision="fp64", compat="physical", cfl=.4, chunk=16,
                 device="cuda:0", keep_results=False, bytes_per_cell=None, max_active_cells=None, backfill=False):
        self.budget = hbm_budget
        # beyond the cells that fill the GPU, admitting more work adds no
        # throughput and only stretches the running requests (processor
        # sharing): an optional cap keeps the excess queued in arrival order,
        # trading median latency for tail latency.  A request larger than the
        # cap still runs, alone.  With backfill, requests smaller than a head
        # of the queue that the cap holds back are admitted past it while
        # they fit under the cap (the large ones keep their arrival order).
        self.max_active_cells = max_active_cells
        self.backfill = backfill
        self.precision, self.compat, self.cfl, self.chunk = precision, compat, cfl, chunk
        self.device = device
        self.keep_results = keep_results
        self.groups = {}
        # state + stage registers + interface states per cell (2D hydro, SSPRK(3,3));
        # both precision modes keep FP64 values in HBM
        self.bytes_per_cell = bytes_per_cell or (4 * 8 * 5 + 2 * 4 * 8 + 8)

    def cost(self, r):
        return r.n * r.n * self.bytes_per_cell

    def _replan_peak(self, extra):
        """Most bytes the plans hold at once if ``extra`` (request costs per
        scheme group) is admitted now: the stale groups re-plan one at a time
        in the loop's order, each new plan built while its old one is alive;
        afterwards, whenever a request finishes, its group re-plans without
        it the same way (a group of one just closes), so room is kept for the
        largest such shrink as well."""
        keys = list(self.groups) + [k for k in extra if k not in self.groups]
        cur = {k: (self.groups[k].held(self.cost) if k in self.groups else 0) for k in keys}
        members = {k: ([self.cost(r) for r in self.groups[k].reqs] if k in self.groups else []) + list(extra.get(k, []))
                   for k in keys}
        total = sum(cur.values())
        peak = total
        for k in keys:
            if (k in self.groups and self.groups[k].stale) or k in extra:
                new = sum(members[k])
                peak = max(peak, total + new)
                total += new - cur[k]
        shrink = max((sum(m) - min(m) for m in members.values() if len(m) > 1), default=0)
        return max(peak, total + shrink)

    def run(self, requests, realtime=True, progress=None):
        """Serve ``requests`` (arrival times relative to the start; with
        ``realtime`` False all are queued at once).  ``progress``: seconds
        between progress lines on stderr (None: silent)."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        finished = []
        rebuilds = 0
        rebuild_s = 0.0
        last = 0.0
        while pending or waiting or any(g.reqs for g in self.groups.values()):
            tn = now()
            if progress is not None and tn - last >= progress:
                last = tn
                _progress_line("served", tn, len(finished), len(requests))
            while pending and (pending[0].arrival <= tn or not realtime):
                waiting.append(pending.popleft())
            extra, admitted = {}, []
            active_cells = sum(x.n * x.n for g in self.groups.values() for x in g.reqs)
            held = None  # the head the cell cap holds back (backfill)
            i = 0
            while i < len(waiting):
                r = waiting[i]
                if held is not None and r.n >= held.n:
                    i += 1
                    continue
                if (self.max_active_cells is not None and active_cells > 0
                        and active_cells + r.n * r.n > self.max_active_cells):
                    if not self.backfill:
                        break
                    held = held or r
                    i += 1
                    continue
                trial = {k: list(v) for k, v in extra.items()}
                trial.setdefault(r.subgrid, []).append(self.cost(r))
                if self._replan_peak(trial) > self.budget:
                    break
                extra = trial
                admitted.append(r)
                del waiting[i]
                active_cells += r.n * r.n
            idle = not admitted and not any(g.reqs or g.stale for g in self.groups.values())
            if waiting and idle and self._replan_peak({waiting[0].subgrid: [self.cost(waiting[0])]}) > self.budget:
                raise ValueError(f"request {waiting[0].rid} ({self.cost(waiting[0])} B) does not fit the HBM "
                                 f"budget of {self.budget} B")
            for r in admitted:
                r.admitted_at = now()
                g = self.groups.get(r.subgrid)
                if g is None:
                    g = self.groups[r.subgrid] = _Group(r.subgrid, self.precision, self.compat, self.cfl, self.device)
                g.reqs.append(r)
                g.stale = True
            for g in self.groups.values():
                if g.stale:
                    r0 = time.perf_counter()
                    g.rebuild()
                    rebuild_s += time.perf_counter() - r0
                    rebuilds += 1
            if not any(g.reqs for g in self.groups.values()):
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - now(), 0.05)))
                continue
            for g in self.groups.values():
                g.advance(self.chunk)
            for g in self.groups.values():
                finished.extend(g.harvest(now, self.keep_results))
        for g in self.groups.values():
            if g.plan:
                g.plan.close()
        out = _summary(finished, now(), rebuilds)
        out["rebuild_s"] = rebuild_s
        return out


def _progress_line(what, t, done, total):
    import sys
    print(f"[serve] {what}: {done}/{total} requests done at {t:.0f} s", file=sys.stderr, flush=True)


def run_fifo(requests, precision="fp64", compat="physical", cfl=.4, chunk=64, realtime=True, keep_results=False,
             device="cuda:0", progress=None):
    """The plain loop: one request at a time in arrival order."""
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0
    done = []
    last = 0.0
    for r in sorted(requests, key=lambda r: r.arrival):
        if progress is not None and now() - last >= progress:
            last = now()
            _progress_line("plain", last, len(done), len(requests))
        if realtime and r.arrival > now():
            time.sleep(r.arrival - now())
        r.admitted_at = now()
        plan = Plan(2, r.subgrid, "hllc", "ssprk(3,3)", [_problem(r, cfl)], precision=precision, varmask=0x17,
                    compat=compat, device=device)
        _initial_state(plan, 0, r)
        plan.reset()
        r.status = plan.run(chunk)[0]
        r.done_at = now()
        if keep_results:
            r.result = plan.state.clone()
        plan.close()
        done.append(r)
    return _summary(done, now(), 0)


def _summary(done, makespan, rebuilds):
    import numpy as np
    lat = np.array([r.latency for r in done]) if done else np.zeros(0)
    zu = sum(r.n * r.n * r.status["steps"] for r in done)
    first = min((r.arrival for r in done), default=0.0)
    return dict(requests=len(done), failed=sum(bool(r.status["failed"]) for r in done), zu=zu,
                seconds=makespan - first, zu_per_s=zu / max(makespan - first, 1e-9),
                p50=float(np.percentile(lat, 50)) if len(lat) else None,
                p99=float(np.percentile(lat, 99)) if len(lat) else None, rebuilds=rebuilds, done=done)
