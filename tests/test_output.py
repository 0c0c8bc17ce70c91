"""The snapshot series sink (output.SnapshotFile) and the per-pass saving
loop of core.run_saving, on the CPU with a scripted stand-in for the plan."""
import numpy as np
import pytest
import torch

from astrea_amd import core
from astrea_amd.output import SnapshotFile


def test_snapshot_file_round_trip(tmp_path):
    path = tmp_path / "run.npz"
    rng = np.random.default_rng(3)
    series = {str(float(t)): rng.standard_normal((5, 4, 8)) for t in (0., .1, .30000000000000004, .5)}
    with SnapshotFile(path) as f:
        for k, a in series.items():
            f(k, a)
        assert f.keys() == list(series)
        with pytest.raises(ValueError):
            f("0.1", series["0.1"])
    with np.load(path) as z:
        assert sorted(z.files, key=float) == list(series)
        for k, a in series.items():
            assert np.array_equal(z[k], a) and z[k].dtype == np.float64
    # appending continues the series
    with SnapshotFile(path, "a") as f:
        f("0.7", np.zeros((5, 4, 8)))
        assert f.keys()[-1] == "0.7"
    with np.load(path) as z:
        assert len(z.files) == 5


class _ScriptedPlan:
    """Stand-in for Plan: the time state follows a scripted list of t values;
    step() advances along it, the state is a constant array tagged by t."""

    def __init__(self, ts, t_end, fail_at=None, max_steps=None):
        self.ts, self.t_end, self.fail_at, self.max_steps = ts, t_end, fail_at, max_steps
        self.k = 0

    def status(self):
        t = self.ts[self.k]
        failed = self.fail_at is not None and self.k >= self.fail_at
        more = t < self.t_end and (self.max_steps is None or self.k < self.max_steps)
        return [dict(t=t, dt=0., steps=self.k, active=more and not failed, failed=failed)]

    def snapshot(self, i, out=None):
        v = torch.full((3, 8), float(self.ts[self.k]), dtype=torch.float64)
        if out is None:
            return v
        out.copy_(v)
        return out

    def step(self, n=1):
        self.k += n


def _saved(plan, t_end, monkeypatch):
    got = []
    # pinned host memory needs a device; the scripted plan lives on the CPU
    real_empty = torch.empty
    monkeypatch.setattr(core.torch, "empty", lambda *a, pin_memory=False, **k: real_empty(*a, **k))
    st = core.run_saving(plan, lambda key, a: got.append((key, a[0, 0])), t_end)
    return got, st


def test_saving_loop_passes(monkeypatch):
    # astrea.py:44-50: a snapshot at t = 0, after every step, and at t_end
    got, st = _saved(_ScriptedPlan([0., .25, .5, .75, 1.], 1.), 1., monkeypatch)
    assert [k for k, _ in got] == ["0.0", "0.25", "0.5", "0.75", "1.0"]
    assert all(float(k) == v for k, v in got) and not st["failed"]


def test_saving_loop_overshoot_and_failure(monkeypatch):
    # t past t_end (the loop condition fails): no snapshot of that state
    got, _ = _saved(_ScriptedPlan([0., .6, 1.0000000000000002], 1.), 1., monkeypatch)
    assert [k for k, _ in got] == ["0.0", "0.6"]
    # a failed step ends the loop without a snapshot of the failed state
    got, st = _saved(_ScriptedPlan([0., .3, .6, .9], 1., fail_at=2), 1., monkeypatch)
    assert [k for k, _ in got] == ["0.0", "0.3"] and st["failed"]
    # a step limit stops after the snapshot of the last state
    got, st = _saved(_ScriptedPlan([0., .1, .2, .3], 1., max_steps=2), 1., monkeypatch)
    assert [k for k, _ in got] == ["0.0", "0.1", "0.2"] and st["steps"] == 2


def test_snapshot_file_groups_and_attributes(tmp_path):
    """The reference's layout (astrea.py:150-186): file attributes, one group
    per run with its attributes and its snapshot series."""
    from astrea_amd.output import read_attrs
    path = tmp_path / "runs.npz"
    with SnapshotFile(path, attrs={"datetime": "202612021200", "seed": 7}) as f:
        f("20261202120000123/0.0", np.ones((4, 8)))
        f("20261202120000123/0.5", np.zeros((4, 8)))
        f.set_attrs("20261202120000123", {"config": "sod", "cells": 4, "elapsed": 1.5})
        with pytest.raises(ValueError):
            f.set_attrs("20261202120000123", {})
        assert f.keys("20261202120000123") == ["0.0", "0.5"] and f.keys() == []
    with SnapshotFile(path, "a") as f:
        f("20261202120001456/0.0", np.ones((4, 8)))
        f.set_attrs("20261202120001456", {"config": "sod", "cells": 4, "elapsed": 2.})
    assert read_attrs(path) == {"datetime": "202612021200", "seed": 7}
    assert read_attrs(path, "20261202120000123")["elapsed"] == 1.5
    with np.load(path) as z:
        assert np.array_equal(z["20261202120000123/0.5"], np.zeros((4, 8)))
        assert sum(k.endswith("/0.0") for k in z.files) == 2
