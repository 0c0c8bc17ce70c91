"""Snapshot series on disk: the output path around astrea.py:44-50.

The reference appends one dataset per pass of its time loop to an HDF5 file,
named str(float(t)), holding the primitive snapshot (convert_conservative of
the grid, transposed by ortho_axis), gzip level 9, with the attribute t; the
datasets of a run sit in a group named by the run's access key, with the run
attributes (astrea.py:165-175, 186), and the file carries datetime and seed
(152-153).  ``Hdf5SnapshotFile`` writes that file itself through the HDF5 C
library (``h5``; h5py is not in this image).  ``SnapshotFile`` writes the
same layout as a zip archive of .npy members, deflate level 9 -- "<t>.npy" or
"<group>/<t>.npy" per snapshot, "attrs.json" / "<group>/attrs.json" per
attribute set -- so ``numpy.load(path)`` opens it as a mapping from
str(float(t)) (or "<group>/" + str(float(t))) to the snapshot.
``open_snapshots`` picks HDF5 for the reference's .hdf5 / .h5 names and the
zip archive otherwise.  Either writes snapshots as they arrive; nothing of
the series is kept in memory.
"""
import json
import os
import posixpath
import zipfile

import numpy as np

from . import h5

HDF5_SIGNATURE = b"\x89HDF\r\n\x1a\n"


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


class Hdf5SnapshotFile:
    """The reference's HDF5 snapshot file (astrea.py:48-50, 151-153,
    166-175, 186), with ``SnapshotFile``'s interface: ``save(key, snapshot)``
    creates the float64 dataset `key` ("<group>/" + str(float(t))), chunked
    and gzip level 9, with the attribute t = float(t); ``set_attrs`` writes
    scalar attributes of the file ("") or of a run's group (str as a
    variable-length UTF-8 string, int as int64, float as float64; a None
    value, such as an unset seed, is left out).  Keys must be new, as for
    h5py's create_dataset."""

    def __init__(self, path, mode="w", attrs=None):
        if mode not in ("w", "a", "x"):
            raise ValueError("mode must be 'w', 'a' or 'x'")
        self.path = path
        self._f = h5.File(path, mode)
        if attrs is not None:
            try:
                self.set_attrs("", attrs)
            except Exception:
                self.close()
                raise

    def __call__(self, key, snapshot):
        key = str(key)
        self._f.create_dataset(key, snapshot, level=9)
        self._f.set_attr(key, "t", float(posixpath.basename(key)))

    def set_attrs(self, group, attrs):
        self._f.require_group(group)
        for k, v in attrs.items():
            if v is not None:
                self._f.set_attr(group, k, v)

    def keys(self, group=""):
        """Snapshot keys str(float(t)) of a group ("" = the top level), by t."""
        out = []
        for n in self._f.datasets(group):
            try:
                float(n)
            except ValueError:
                continue
            out.append(n)
        return sorted(out, key=float)

    def close(self):
        if self._f is not None:
            self._f.close()
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def open_snapshots(path, mode="w", attrs=None):
    """The snapshot sink for `path`: the HDF5 file for .hdf5 / .h5 names
    (an error if libhdf5 cannot be loaded), the zip archive otherwise."""
    if os.path.splitext(str(path))[1].lower() in (".hdf5", ".h5"):
        return Hdf5SnapshotFile(path, mode, attrs)
    return SnapshotFile(path, mode, attrs)


def _is_hdf5(path):
    with open(path, "rb") as f:
        return f.read(8) == HDF5_SIGNATURE


def read_attrs(path, group=""):
    """The attributes ``set_attrs`` wrote for `group` (either format)."""
    if _is_hdf5(path):
        with h5.File(path, "r") as f:
            return {k: f.read_attr(group, k) for k in f.attr_names(group)}
    with zipfile.ZipFile(path) as zf:
        return json.loads(zf.read(posixpath.join(group, "attrs.json") if group else "attrs.json"))


def snapshot_keys(path, group=""):
    """Snapshot keys str(float(t)) of `group` in either format, by t."""
    if _is_hdf5(path):
        with h5.File(path, "r") as f:
            names = f.datasets(group)
    else:
        with zipfile.ZipFile(path) as zf:
            names = [posixpath.basename(n)[:-4] for n in zf.namelist()
                     if n.endswith(".npy") and posixpath.dirname(n) == group]
    out = []
    for n in names:
        try:
            float(n)
        except ValueError:
            continue
        out.append(n)
    return sorted(out, key=float)


def read_snapshot(path, key):
    """The snapshot `key` ("<group>/" + str(float(t))) of either format."""
    if _is_hdf5(path):
        with h5.File(path, "r") as f:
            return f.read_dataset(str(key))
    with np.load(path, allow_pickle=False) as z:
        return z[str(key)]
