"""The HDF5 snapshot file of the reference, written through the HDF5 C
library (libhdf5, bound with ctypes; h5py is not in this image).

The reference opens its file with h5py (astrea.py:48-50, 151-153, 166-175,
185-186).  What h5py stores for the values it writes there is reproduced:

* a snapshot: a float64 dataset named str(float(t)) in the run's group,
  chunked, gzip (deflate) level 9, with the scalar float64 attribute ``t``;
* a Python ``str`` attribute: a scalar variable-length UTF-8 string;
* an ``int``: a scalar int64; a ``float``: a scalar float64.

The library is looked up as ``libhdf5`` on the loader path, then in
``$ASTREA_HDF5_LIB`` and the conda prefix of this image (/opt/conda/lib).
"""
import ctypes
import ctypes.util
import os

import numpy as np

hid_t = ctypes.c_int64
herr_t = ctypes.c_int
htri_t = ctypes.c_int
hsize_t = ctypes.c_ulonglong

# H5Fpublic.h, H5Spublic.h, H5Ppublic.h, H5Tpublic.h, H5public.h
ACC_RDONLY, ACC_RDWR, ACC_TRUNC, ACC_EXCL = 0x0, 0x1, 0x2, 0x4
P_DEFAULT = 0
S_ALL = 0
S_SCALAR = 0
E_DEFAULT = 0
T_VARIABLE = ctypes.c_size_t(-1).value
T_CSET_UTF8 = 1
INDEX_NAME = 0
ITER_INC = 0
# h5py's chunk size bounds (bytes) for an auto-chunked dataset
CHUNK_MAX = 1 << 20


class H5GInfo(ctypes.Structure):
    _fields_ = [("storage_type", ctypes.c_int), ("nlinks", hsize_t), ("max_corder", ctypes.c_int64),
                ("mounted", ctypes.c_bool)]


_lib = None


def _candidates():
    env = os.environ.get("ASTREA_HDF5_LIB")
    if env:
        yield env
    found = ctypes.util.find_library("hdf5")
    if found:
        yield found
    for name in ("libhdf5.so", "libhdf5.so.103"):
        yield os.path.join("/opt/conda/lib", name)


def available():
    try:
        lib()
        return True
    except OSError:
        return False


def lib():
    """The loaded libhdf5 with the prototypes used here (OSError if none)."""
    global _lib
    if _lib is not None:
        return _lib
    err = None
    for c in _candidates():
        try:
            h = ctypes.CDLL(c)
            break
        except OSError as e:
            err = e
    else:
        raise OSError(f"libhdf5 not found (set ASTREA_HDF5_LIB): {err}")
    proto = {
        "H5open": (herr_t, []),
        "H5Eset_auto2": (herr_t, [hid_t, ctypes.c_void_p, ctypes.c_void_p]),
        "H5Fcreate": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t]),
        "H5Fopen": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t]),
        "H5Fclose": (herr_t, [hid_t]),
        "H5Gcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t]),
        "H5Gopen2": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Gclose": (herr_t, [hid_t]),
        "H5Gget_info": (herr_t, [hid_t, ctypes.POINTER(H5GInfo)]),
        "H5Lexists": (htri_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Lget_name_by_idx": (ctypes.c_ssize_t, [hid_t, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, hsize_t,
                                                  ctypes.c_char_p, ctypes.c_size_t, hid_t]),
        "H5Screate": (hid_t, [ctypes.c_int]),
        "H5Screate_simple": (hid_t, [ctypes.c_int, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]),
        "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid_t]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [hid_t, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]),
        "H5Sclose": (herr_t, [hid_t]),
        "H5Pcreate": (hid_t, [hid_t]),
        "H5Pset_chunk": (herr_t, [hid_t, ctypes.c_int, ctypes.POINTER(hsize_t)]),
        "H5Pset_deflate": (herr_t, [hid_t, ctypes.c_uint]),
        "H5Pclose": (herr_t, [hid_t]),
        "H5Tcopy": (hid_t, [hid_t]),
        "H5Tset_size": (herr_t, [hid_t, ctypes.c_size_t]),
        "H5Tset_cset": (herr_t, [hid_t, ctypes.c_int]),
        "H5Tclose": (herr_t, [hid_t]),
        "H5Dcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
        "H5Dopen2": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Dwrite": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Dread": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Dget_space": (hid_t, [hid_t]),
        "H5Dclose": (herr_t, [hid_t]),
        "H5Acreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t]),
        "H5Aopen": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Aexists": (htri_t, [hid_t, ctypes.c_char_p]),
        "H5Adelete": (herr_t, [hid_t, ctypes.c_char_p]),
        "H5Awrite": (herr_t, [hid_t, hid_t, ctypes.c_void_p]),
        "H5Aread": (herr_t, [hid_t, hid_t, ctypes.c_void_p]),
        "H5Aget_type": (hid_t, [hid_t]),
        "H5Aget_num_attrs": (ctypes.c_int, [hid_t]),
        "H5Aget_name_by_idx": (ctypes.c_ssize_t, [hid_t, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, hsize_t,
                                                  ctypes.c_char_p, ctypes.c_size_t, hid_t]),
        "H5Aclose": (herr_t, [hid_t]),
        "H5Tget_class": (ctypes.c_int, [hid_t]),
        "H5Tis_variable_str": (htri_t, [hid_t]),
        "H5free_memory": (herr_t, [ctypes.c_void_p]),
    }
    for name, (res, args) in proto.items():
        f = getattr(h, name)
        f.restype, f.argtypes = res, args
    if h.H5open() < 0:
        raise OSError("H5open failed")
    # errors come back as return codes, raised here with the call's name
    h.H5Eset_auto2(E_DEFAULT, None, None)
    h.T_NATIVE_DOUBLE = hid_t.in_dll(h, "H5T_NATIVE_DOUBLE_g").value
    h.T_NATIVE_INT64 = hid_t.in_dll(h, "H5T_NATIVE_INT64_g").value
    h.T_C_S1 = hid_t.in_dll(h, "H5T_C_S1_g").value
    h.P_DATASET_CREATE = hid_t.in_dll(h, "H5P_CLS_DATASET_CREATE_ID_g").value
    _lib = h
    return h


def _ok(rc, what):
    if rc < 0:
        raise RuntimeError(f"HDF5: {what} failed")
    return rc


# H5T_class_t values read back by read_attr
T_INTEGER, T_FLOAT, T_STRING = 0, 1, 3


def guess_chunk(shape, itemsize):
    """Chunk shape of an auto-chunked dataset: the whole array, halving the
    largest extent until a chunk holds at most CHUNK_MAX bytes (the chunking
    changes how the file stores the data, not its values)."""
    c = [max(int(s), 1) for s in shape]
    while int(np.prod(c)) * itemsize > CHUNK_MAX and max(c) > 1:
        k = int(np.argmax(c))
        c[k] = (c[k] + 1) // 2
    return c


class File:
    """An open HDF5 file: groups, gzip-9 float64 datasets and scalar
    attributes as h5py writes them for the reference's values."""

    def __init__(self, path, mode="a"):
        h = lib()
        self.h = h
        self.path = path
        p = os.fsencode(path)
        if mode == "r":
            self.fid = _ok(h.H5Fopen(p, ACC_RDONLY, P_DEFAULT), f"open {path}")
        elif mode == "a" and os.path.exists(path):
            self.fid = _ok(h.H5Fopen(p, ACC_RDWR, P_DEFAULT), f"open {path}")
        elif mode in ("w", "a", "x"):
            flag = ACC_EXCL if mode == "x" else ACC_TRUNC
            self.fid = _ok(h.H5Fcreate(p, flag, P_DEFAULT, P_DEFAULT), f"create {path}")
        else:
            raise ValueError("mode must be 'r', 'w', 'a' or 'x'")

    def exists(self, name):
        # every intermediate link must exist before the last one is asked for
        parts = name.strip("/").split("/")
        for i in range(len(parts)):
            sub = "/".join(parts[:i + 1]).encode()
            if self.h.H5Lexists(self.fid, sub, P_DEFAULT) <= 0:
                return False
        return True

    def require_group(self, name):
        if not name:
            return
        parts = name.strip("/").split("/")
        for i in range(len(parts)):
            sub = "/".join(parts[:i + 1])
            if not self.exists(sub):
                g = _ok(self.h.H5Gcreate2(self.fid, sub.encode(), P_DEFAULT, P_DEFAULT, P_DEFAULT),
                        f"create group {sub}")
                self.h.H5Gclose(g)

    def _open_obj(self, name):
        """(id, close) of the group or dataset `name` ("" = the root group)."""
        h = self.h
        if not name or name == "/":
            return _ok(h.H5Gopen2(self.fid, b"/", P_DEFAULT), "open /"), h.H5Gclose
        g = h.H5Gopen2(self.fid, name.encode(), P_DEFAULT)
        if g >= 0:
            return g, h.H5Gclose
        return _ok(h.H5Dopen2(self.fid, name.encode(), P_DEFAULT), f"open {name}"), h.H5Dclose

    def create_dataset(self, name, data, level=9):
        """A new chunked, deflate-`level` dataset (h5py's create_dataset with
        compression="gzip", compression_opts=level); an existing name is an
        error, as it is for h5py."""
        h = self.h
        if self.exists(name):
            raise ValueError(f"{name!r} already exists in {self.path}")
        parent = name.rsplit("/", 1)[0] if "/" in name else ""
        self.require_group(parent)
        a = np.ascontiguousarray(data, dtype=np.float64)
        dims = (hsize_t * max(a.ndim, 1))(*(a.shape or (1,)))
        space = _ok(h.H5Screate_simple(a.ndim, dims, None), "dataspace") if a.ndim else _ok(h.H5Screate(S_SCALAR),
                                                                                                  "dataspace")
        dcpl = _ok(h.H5Pcreate(h.P_DATASET_CREATE), "dataset properties")
        try:
            if a.ndim:
                ch = guess_chunk(a.shape, a.itemsize)
                _ok(h.H5Pset_chunk(dcpl, a.ndim, (hsize_t * a.ndim)(*ch)), "chunk")
                _ok(h.H5Pset_deflate(dcpl, level), "deflate")
            d = _ok(h.H5Dcreate2(self.fid, name.encode(), h.T_NATIVE_DOUBLE, space, P_DEFAULT, dcpl, P_DEFAULT),
                    f"create dataset {name}")
            try:
                _ok(h.H5Dwrite(d, h.T_NATIVE_DOUBLE, S_ALL, S_ALL, P_DEFAULT, a.ctypes.data), f"write {name}")
            finally:
                h.H5Dclose(d)
        finally:
            h.H5Pclose(dcpl)
            h.H5Sclose(space)

    def read_dataset(self, name):
        h = self.h
        d = _ok(h.H5Dopen2(self.fid, name.encode(), P_DEFAULT), f"open {name}")
        try:
            space = _ok(h.H5Dget_space(d), "dataspace")
            n = h.H5Sget_simple_extent_ndims(space)
            dims = (hsize_t * max(n, 1))()
            h.H5Sget_simple_extent_dims(space, dims, None)
            h.H5Sclose(space)
            out = np.empty(tuple(dims[:n]), dtype=np.float64)
            _ok(h.H5Dread(d, h.T_NATIVE_DOUBLE, S_ALL, S_ALL, P_DEFAULT, out.ctypes.data), f"read {name}")
            return out
        finally:
            h.H5Dclose(d)

    def set_attr(self, obj, key, value):
        """A scalar attribute of the group or dataset `obj`, replacing an
        existing one of that name (h5py's attrs[key] = value)."""
        h = self.h
        if isinstance(value, (bool, np.bool_)) or value is None:
            raise TypeError(f"attribute {key!r}: {type(value).__name__} is not written")
        oid, close = self._open_obj(obj)
        space = typ = aid = -1
        try:
            if h.H5Aexists(oid, key.encode()) > 0:
                _ok(h.H5Adelete(oid, key.encode()), f"replace attribute {key}")
            space = _ok(h.H5Screate(S_SCALAR), "dataspace")
            if isinstance(value, str):
                typ = _ok(h.H5Tcopy(h.T_C_S1), "string type")
                _ok(h.H5Tset_size(typ, T_VARIABLE), "string size")
                _ok(h.H5Tset_cset(typ, T_CSET_UTF8), "string charset")
                raw = ctypes.c_char_p(value.encode("utf-8"))
                buf, mem = ctypes.byref(raw), typ
            elif isinstance(value, (int, np.integer)):
                raw = ctypes.c_int64(int(value))
                buf, mem = ctypes.byref(raw), h.T_NATIVE_INT64
            elif isinstance(value, (float, np.floating)):
                raw = ctypes.c_double(float(value))
                buf, mem = ctypes.byref(raw), h.T_NATIVE_DOUBLE
            else:
                raise TypeError(f"attribute {key!r}: {type(value).__name__} is not written")
            aid = _ok(h.H5Acreate2(oid, key.encode(), mem, space, P_DEFAULT, P_DEFAULT), f"create attribute {key}")
            _ok(h.H5Awrite(aid, mem, buf), f"write attribute {key}")
        finally:
            if aid >= 0:
                h.H5Aclose(aid)
            if typ >= 0:
                h.H5Tclose(typ)
            if space >= 0:
                h.H5Sclose(space)
            close(oid)

    def read_attr(self, obj, key):
        h = self.h
        oid, close = self._open_obj(obj)
        aid = typ = -1
        try:
            aid = _ok(h.H5Aopen(oid, key.encode(), P_DEFAULT), f"open attribute {key}")
            typ = _ok(h.H5Aget_type(aid), "attribute type")
            cls = h.H5Tget_class(typ)
            if cls == T_STRING and h.H5Tis_variable_str(typ) > 0:
                ptr = ctypes.c_char_p()
                _ok(h.H5Aread(aid, typ, ctypes.byref(ptr)), f"read attribute {key}")
                out = ptr.value.decode("utf-8")
                h.H5free_memory(ptr)
                return out
            if cls == T_INTEGER:
                v = ctypes.c_int64()
                _ok(h.H5Aread(aid, h.T_NATIVE_INT64, ctypes.byref(v)), f"read attribute {key}")
                return int(v.value)
            if cls == T_FLOAT:
                v = ctypes.c_double()
                _ok(h.H5Aread(aid, h.T_NATIVE_DOUBLE, ctypes.byref(v)), f"read attribute {key}")
                return float(v.value)
            raise TypeError(f"attribute {key!r}: HDF5 type class {cls} is not read")
        finally:
            if typ >= 0:
                h.H5Tclose(typ)
            if aid >= 0:
                h.H5Aclose(aid)
            close(oid)

    def attr_names(self, obj=""):
        """Attribute names of the group or dataset `obj`, in name order."""
        h = self.h
        oid, close = self._open_obj(obj)
        try:
            out = []
            for i in range(_ok(h.H5Aget_num_attrs(oid), "attribute count")):
                n = _ok(h.H5Aget_name_by_idx(oid, b".", INDEX_NAME, ITER_INC, i, None, 0, P_DEFAULT), "attribute name")
                buf = ctypes.create_string_buffer(n + 1)
                h.H5Aget_name_by_idx(oid, b".", INDEX_NAME, ITER_INC, i, buf, n + 1, P_DEFAULT)
                out.append(buf.value.decode())
            return out
        finally:
            close(oid)

    def names(self, group=""):
        """Link names in `group` ("" = the root), in name order."""
        h = self.h
        gid = _ok(h.H5Gopen2(self.fid, (group or "/").encode(), P_DEFAULT), f"open {group or '/'}")
        try:
            info = H5GInfo()
            _ok(h.H5Gget_info(gid, ctypes.byref(info)), "group info")
            out = []
            for i in range(info.nlinks):
                n = _ok(h.H5Lget_name_by_idx(gid, b".", INDEX_NAME, ITER_INC, i, None, 0, P_DEFAULT), "link name")
                buf = ctypes.create_string_buffer(n + 1)
                h.H5Lget_name_by_idx(gid, b".", INDEX_NAME, ITER_INC, i, buf, n + 1, P_DEFAULT)
                out.append(buf.value.decode())
            return out
        finally:
            h.H5Gclose(gid)

    def datasets(self, group=""):
        """Names of the datasets (not groups) in `group`, in name order."""
        out = []
        for n in self.names(group):
            path = f"{group}/{n}" if group else n
            d = self.h.H5Dopen2(self.fid, path.encode(), P_DEFAULT)
            if d >= 0:
                self.h.H5Dclose(d)
                out.append(n)
        return out

    def close(self):
        if self.fid is not None:
            _ok(self.h.H5Fclose(self.fid), f"close {self.path}")
            self.fid = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False
