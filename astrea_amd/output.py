"""Snapshot series on disk: the output path around astrea.py:44-50.

The reference appends one dataset per pass of its time loop to an HDF5 file,
named str(float(t)), holding the primitive snapshot (convert_conservative of
the grid, transposed by ortho_axis), gzip level 9, with the attribute t; the
datasets of a run sit in a group named by the run's access key, with the run
attributes (astrea.py:165-175, 186), and the file carries datetime and seed
(152-153).  h5py is not a dependency here: ``SnapshotFile`` writes the same
layout as a zip archive of .npy members, deflate level 9 -- "<t>.npy" or
"<group>/<t>.npy" per snapshot, "attrs.json" / "<group>/attrs.json" per
attribute set -- so ``numpy.load(path)`` opens it as a mapping from
str(float(t)) (or "<group>/" + str(float(t))) to the snapshot.  Members are
written as they arrive; nothing of the series is kept in memory.
"""
import json
import posixpath
import zipfile

import numpy as np


class SnapshotFile:
    """Callable sink ``save(key, snapshot)`` for ``core.core_run(..., save=)``.

    Keys must be new (the reference's create_dataset refuses an existing
    name); opening an existing file with mode "a" appends to its series."""

    def __init__(self, path, mode="w", attrs=None):
        if mode not in ("w", "a", "x"):
            raise ValueError("mode must be 'w', 'a' or 'x'")
        self.path = path
        self._zf = zipfile.ZipFile(path, mode, compression=zipfile.ZIP_DEFLATED, compresslevel=9, allowZip64=True)
        self._names = set(self._zf.namelist())
        if attrs is not None:
            self.set_attrs("", attrs)

    def _new(self, name):
        if name in self._names:
            raise ValueError(f"{name!r} already exists in {self.path}")
        self._names.add(name)
        return name

    def __call__(self, key, snapshot):
        name = self._new(str(key) + ".npy")
        a = np.ascontiguousarray(snapshot)
        with self._zf.open(name, "w", force_zip64=True) as f:
            np.lib.format.write_array(f, a, allow_pickle=False)

    def set_attrs(self, group, attrs):
        """Attributes of the file (group "") or of a run's group, written once."""
        name = self._new(posixpath.join(group, "attrs.json") if group else "attrs.json")
        self._zf.writestr(name, json.dumps(attrs, sort_keys=True))

    def keys(self, group=""):
        """Snapshot keys str(float(t)) of a group ("" = the top level), by t."""
        out = [posixpath.basename(n)[:-4] for n in self._names
               if n.endswith(".npy") and posixpath.dirname(n) == group]
        return sorted(out, key=float)

    def close(self):
        if self._zf is not None:
            self._zf.close()
            self._zf = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def read_attrs(path, group=""):
    """The attributes ``SnapshotFile.set_attrs`` wrote for `group`."""
    with zipfile.ZipFile(path) as zf:
        return json.loads(zf.read(posixpath.join(group, "attrs.json") if group else "attrs.json"))
