"""The snapshot series sinks (the reference's HDF5 file through libhdf5,
output.Hdf5SnapshotFile, and the zip archive, output.SnapshotFile) and the
per-pass saving loop of core.run_saving, on the CPU with a scripted stand-in
for the plan."""
import os

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


def _h5dump():
    import shutil
    return shutil.which("h5dump") or next((p for p in ("/opt/conda/bin/h5dump",) if os.path.exists(p)), None)


def test_hdf5_snapshot_file_layout(tmp_path):
    """The reference's HDF5 file (astrea.py:48-50, 151-153, 166-175, 186)
    written through libhdf5: file attributes, a group per run with its
    attributes (an attribute set again is replaced, as astrea.py:186 does for
    'elapsed'), float64 snapshots named str(float(t)) with the attribute t,
    appended runs; an existing snapshot name is refused, as h5py's
    create_dataset refuses it."""
    from astrea_amd import h5
    from astrea_amd.output import Hdf5SnapshotFile, open_snapshots, read_attrs, read_snapshot, snapshot_keys
    if not h5.available():
        pytest.skip("libhdf5 cannot be loaded")
    path = tmp_path / "runs.hdf5"
    rng = np.random.default_rng(3)
    a0, a1, a2 = rng.random((16, 8)), rng.random((300, 200, 8)), rng.random((5, 7, 8))
    with open_snapshots(path, attrs={"datetime": "202612031200", "seed": 7}) as f:
        assert isinstance(f, Hdf5SnapshotFile)
        f("20261203120000123/0.0", a0)
        f("20261203120000123/0.125", a1)
        with pytest.raises(ValueError):
            f("20261203120000123/0.125", a1)
        f.set_attrs("20261203120000123", {"config": "sod", "cells": 16, "cfl": .5, "subgrid": "ppm"})
        f.set_attrs("20261203120000123", {"elapsed": 1.5})
        f.set_attrs("20261203120000123", {"elapsed": 2.5})
        assert f.keys("20261203120000123") == ["0.0", "0.125"] and f.keys() == []
    with open_snapshots(path, "a") as f:
        f("20261203120001456/0.0", a2)
        f.set_attrs("20261203120001456", {"config": "sod", "cells": 16, "seed": None})
    assert read_attrs(path) == {"datetime": "202612031200", "seed": 7}
    assert read_attrs(path, "20261203120000123") == {"config": "sod", "cells": 16, "cfl": .5, "subgrid": "ppm",
                                                     "elapsed": 2.5}
    assert read_attrs(path, "20261203120001456") == {"config": "sod", "cells": 16}
    assert snapshot_keys(path, "20261203120000123") == ["0.0", "0.125"]
    for k, a in (("20261203120000123/0.0", a0), ("20261203120000123/0.125", a1), ("20261203120001456/0.0", a2)):
        assert np.array_equal(read_snapshot(path, k), a)
    with h5.File(str(path), "r") as f:
        assert f.read_attr("20261203120000123/0.125", "t") == 0.125
        assert sorted(f.names()) == ["20261203120000123", "20261203120001456"]
    # the format as the HDF5 tools see it: variable-length UTF-8 strings,
    # 64-bit integers and floats, deflate level 9 on chunked datasets
    tool = _h5dump()
    if tool is None:
        return
    import subprocess
    out = subprocess.run([tool, "-H", "-p", str(path)], capture_output=True, text=True, check=True).stdout
    assert "STRSIZE H5T_VARIABLE;" in out and "CSET H5T_CSET_UTF8;" in out
    assert 'ATTRIBUTE "seed" {\n      DATATYPE  H5T_STD_I64LE' in out
    assert 'ATTRIBUTE "cfl" {\n         DATATYPE  H5T_IEEE_F64LE' in out
    assert out.count("COMPRESSION DEFLATE { LEVEL 9 }") == 3
    assert "DATASPACE  SIMPLE { ( 300, 200, 8 ) / ( 300, 200, 8 ) }" in out
    assert out.count('ATTRIBUTE "t" {') == 3


def test_zip_archive_for_other_names(tmp_path):
    """Names other than .hdf5 / .h5 keep the zip archive of .npy members."""
    from astrea_amd.output import open_snapshots, read_snapshot, snapshot_keys
    path = tmp_path / "runs.npz"
    with open_snapshots(path, attrs={"seed": 1}) as f:
        assert isinstance(f, SnapshotFile)
        f("g/0.5", np.ones((3, 8)))
    assert snapshot_keys(path, "g") == ["0.5"] and np.array_equal(read_snapshot(path, "g/0.5"), np.ones((3, 8)))


def test_hdf5_chunks_and_refused_values(tmp_path):
    """Chunks of an auto-chunked snapshot hold at most 1 MiB and tile the
    array's shape; attribute values the reference never writes (bool, None
    as a single value, containers) are refused rather than stored in some
    other HDF5 type."""
    from astrea_amd import h5
    for shape in ((1024, 1024, 8), (64, 8), (3, 5, 7), (40000, 8)):
        c = h5.guess_chunk(shape, 8)
        assert len(c) == len(shape) and all(1 <= a <= b for a, b in zip(c, shape))
        assert int(np.prod(c)) * 8 <= h5.CHUNK_MAX or int(np.prod(c)) == 1
    if not h5.available():
        pytest.skip("libhdf5 cannot be loaded")
    with h5.File(str(tmp_path / "a.hdf5"), "w") as f:
        f.require_group("g")
        for bad in (True, None, [1, 2], {"a": 1}):
            with pytest.raises(TypeError):
                f.set_attr("g", "x", bad)
        f.set_attr("g", "n", np.int64(3))
        f.set_attr("g", "v", np.float32(.5))
        assert f.read_attr("g", "n") == 3 and f.read_attr("g", "v") == .5
