"""Snapshot series on disk: the output path around astrea.py:44-50.

The reference appends one dataset per pass of its time loop to an HDF5 file,
named str(float(t)), holding the primitive snapshot (convert_conservative of
the grid, transposed by ortho_axis), gzip level 9, with the attribute t.
h5py is not a dependency here: ``SnapshotFile`` writes the same series as a
zip archive of .npy members, deflate level 9, so ``numpy.load(path)`` opens it
as a mapping from str(float(t)) to the snapshot and float(key) gives t back.
Members are written as they arrive; nothing of the series is kept in memory.
"""
import zipfile

import numpy as np


class SnapshotFile:
    """Callable sink ``save(key, snapshot)`` for ``core.core_run(..., save=)``.

    Keys must be new (the reference's create_dataset refuses an existing
    name); opening an existing file with mode "a" appends to its series."""

    def __init__(self, path, mode="w"):
        if mode not in ("w", "a", "x"):
            raise ValueError("mode must be 'w', 'a' or 'x'")
        self.path = path
        self._zf = zipfile.ZipFile(path, mode, compression=zipfile.ZIP_DEFLATED, compresslevel=9, allowZip64=True)
        self._names = set(self._zf.namelist())

    def __call__(self, key, snapshot):
        name = str(key) + ".npy"
        if name in self._names:
            raise ValueError(f"snapshot {key!r} already exists in {self.path}")
        a = np.ascontiguousarray(snapshot)
        with self._zf.open(name, "w", force_zip64=True) as f:
            np.lib.format.write_array(f, a, allow_pickle=False)
        self._names.add(name)

    def keys(self):
        return sorted((n[:-4] for n in self._names), key=float)

    def close(self):
        if self._zf is not None:
            self._zf.close()
            self._zf = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False
