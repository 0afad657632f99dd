"""HDF5 archive for solver checkpoints, over the HDF5 C library through ctypes (h5py is not part of
this environment).

Mirrors how the reference persists its solver, with TRIQS's HDFArchive and register_class
(python/triqs_xca/block_sparse_solver.py:1174-1204, 1441-1443; test/python/h5_io_bs.py:30-52):

    with HDFArchive("run.h5", "w") as A:
        A["S"] = S              # an object with __reduce_to_dict__ is stored as a group
    with HDFArchive("run.h5", "r") as A:
        S2 = A["S"]             # rebuilt by the registered class's __factory_from_dict__

Layout, following TRIQS's h5 conventions: a dict or a registered object is a group (an object's
group carries the string attribute "Format" with its class name); a list or tuple is a group with
keys "0", "1", ... and Format "list" or "tuple"; a numpy array or scalar is a dataset, a complex one
stored as real with a trailing dimension of 2 and the attribute "__complex__" = 1; a string is a
fixed-length string dataset; None is a group with Format "None".
"""
import ctypes as C
import os

import numpy as np

_LIB_CANDIDATES = [os.environ.get("XCA_HDF5_LIB", ""), "libhdf5.so", "/opt/conda/lib/libhdf5.so"]

hid_t = C.c_int64
herr_t = C.c_int
hsize_t = C.c_ulonglong

H5F_ACC_RDONLY, H5F_ACC_TRUNC = 0x0000, 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3
H5I_GROUP, H5I_DATASET = 2, 5
H5_INDEX_NAME, H5_ITER_INC = 0, 0

_h5 = None
_registry = {}


def register_class(cls):
    """Objects of cls are stored through cls.__reduce_to_dict__ and read back through
    cls.__factory_from_dict__(name, dict), as with TRIQS's h5.formats.register_class."""
    _registry[cls.__name__] = cls


class H5Error(RuntimeError):
    pass


def lib():
    global _h5
    if _h5 is None:
        err = None
        for path in _LIB_CANDIDATES:
            if not path:
                continue
            try:
                L = C.CDLL(path)
                break
            except OSError as e:
                err = e
        else:
            raise H5Error(f"the HDF5 C library is not available ({err}); set XCA_HDF5_LIB to its path")
        sig = {
            "H5open": (herr_t, []), "H5Eset_auto2": (herr_t, [hid_t, C.c_void_p, C.c_void_p]),
            "H5Fcreate": (hid_t, [C.c_char_p, C.c_uint, hid_t, hid_t]), "H5Fopen": (hid_t, [C.c_char_p, C.c_uint, hid_t]),
            "H5Fclose": (herr_t, [hid_t]),
            "H5Gcreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t]), "H5Gopen2": (hid_t, [hid_t, C.c_char_p, hid_t]),
            "H5Gclose": (herr_t, [hid_t]),
            "H5Screate": (hid_t, [C.c_int]), "H5Screate_simple": (hid_t, [C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
            "H5Sclose": (herr_t, [hid_t]), "H5Sget_simple_extent_ndims": (C.c_int, [hid_t]),
            "H5Sget_simple_extent_dims": (C.c_int, [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
            "H5Dcreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
            "H5Dopen2": (hid_t, [hid_t, C.c_char_p, hid_t]),
            "H5Dwrite": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
            "H5Dread": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
            "H5Dget_space": (hid_t, [hid_t]), "H5Dget_type": (hid_t, [hid_t]), "H5Dclose": (herr_t, [hid_t]),
            "H5Tcopy": (hid_t, [hid_t]), "H5Tset_size": (herr_t, [hid_t, C.c_size_t]), "H5Tget_size": (C.c_size_t, [hid_t]),
            "H5Tget_class": (C.c_int, [hid_t]), "H5Tclose": (herr_t, [hid_t]),
            "H5Acreate2": (hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t]), "H5Aopen": (hid_t, [hid_t, C.c_char_p, hid_t]),
            "H5Awrite": (herr_t, [hid_t, hid_t, C.c_void_p]), "H5Aread": (herr_t, [hid_t, hid_t, C.c_void_p]),
            "H5Aexists": (C.c_int, [hid_t, C.c_char_p]), "H5Aget_type": (hid_t, [hid_t]), "H5Aclose": (herr_t, [hid_t]),
            "H5Oopen": (hid_t, [hid_t, C.c_char_p, hid_t]), "H5Oclose": (herr_t, [hid_t]), "H5Iget_type": (C.c_int, [hid_t]),
            "H5Lget_name_by_idx": (C.c_ssize_t, [hid_t, C.c_char_p, C.c_int, C.c_int, hsize_t, C.c_char_p, C.c_size_t, hid_t]),
            "H5Lexists": (C.c_int, [hid_t, C.c_char_p, hid_t]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        if L.H5open() < 0:
            raise H5Error("H5open failed")
        L.H5Eset_auto2(0, None, None)  # errors are reported through return codes here
        _h5 = L
    return _h5


def _t(name):
    return hid_t.in_dll(lib(), name).value


def _ok(v, what):
    if v < 0:
        raise H5Error(f"HDF5: {what} failed")
    return v


def _nlinks(g):
    n = 0
    while lib().H5Lget_name_by_idx(g, b".", H5_INDEX_NAME, H5_ITER_INC, n, None, 0, H5P_DEFAULT) >= 0:
        n += 1
    return n


def _link_names(g):
    names = []
    for i in range(_nlinks(g)):
        size = lib().H5Lget_name_by_idx(g, b".", H5_INDEX_NAME, H5_ITER_INC, i, None, 0, H5P_DEFAULT)
        buf = C.create_string_buffer(size + 1)
        lib().H5Lget_name_by_idx(g, b".", H5_INDEX_NAME, H5_ITER_INC, i, buf, size + 1, H5P_DEFAULT)
        names.append(buf.value.decode())
    return names


def _write_attr_str(obj, name, value):
    L = lib()
    data = value.encode()
    t = _ok(L.H5Tcopy(_t("H5T_C_S1_g")), "H5Tcopy")
    L.H5Tset_size(t, max(1, len(data)))
    sp = L.H5Screate(H5S_SCALAR)
    a = _ok(L.H5Acreate2(obj, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT), f"attribute {name}")
    _ok(L.H5Awrite(a, t, C.create_string_buffer(data, max(1, len(data)))), f"attribute {name}")
    L.H5Aclose(a), L.H5Sclose(sp), L.H5Tclose(t)


def _read_attr_str(obj, name):
    L = lib()
    if L.H5Aexists(obj, name.encode()) <= 0:
        return None
    a = L.H5Aopen(obj, name.encode(), H5P_DEFAULT)
    t = L.H5Aget_type(a)
    size = L.H5Tget_size(t)
    buf = C.create_string_buffer(size + 1)
    L.H5Aread(a, t, buf)
    L.H5Tclose(t), L.H5Aclose(a)
    return buf.value.decode()


def _write_attr_int(obj, name, value):
    L = lib()
    sp = L.H5Screate(H5S_SCALAR)
    tt = _t("H5T_NATIVE_LLONG_g")
    a = _ok(L.H5Acreate2(obj, name.encode(), tt, sp, H5P_DEFAULT, H5P_DEFAULT), f"attribute {name}")
    v = C.c_longlong(value)
    L.H5Awrite(a, tt, C.byref(v))
    L.H5Aclose(a), L.H5Sclose(sp)


def _write_dataset(g, key, value):
    L = lib()
    if isinstance(value, str):
        data = value.encode()
        t = L.H5Tcopy(_t("H5T_C_S1_g"))
        L.H5Tset_size(t, max(1, len(data)))
        sp = L.H5Screate(H5S_SCALAR)
        d = _ok(L.H5Dcreate2(g, key.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"dataset {key}")
        _ok(L.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, C.create_string_buffer(data, max(1, len(data)))), f"write {key}")
        L.H5Dclose(d), L.H5Sclose(sp), L.H5Tclose(t)
        return
    a = np.asarray(value)
    cplx = np.iscomplexobj(a)
    # np.array, not np.ascontiguousarray, which turns a 0-d array (a scalar) into a 1-d one
    if cplx:
        a = np.array(np.stack([a.real, a.imag], axis=-1), dtype=np.float64, order="C")
    elif a.dtype == np.bool_ or np.issubdtype(a.dtype, np.integer):
        a = np.array(a, dtype=np.int64, order="C")
    else:
        a = np.array(a, dtype=np.float64, order="C")
    tt = _t("H5T_NATIVE_LLONG_g") if a.dtype == np.int64 else _t("H5T_NATIVE_DOUBLE_g")
    if a.ndim == 0:
        sp = L.H5Screate(H5S_SCALAR)
    else:
        dims = (hsize_t * a.ndim)(*a.shape)
        sp = L.H5Screate_simple(a.ndim, dims, None)
    d = _ok(L.H5Dcreate2(g, key.encode(), tt, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"dataset {key}")
    if a.size:
        _ok(L.H5Dwrite(d, tt, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data_as(C.c_void_p)), f"write {key}")
    if cplx:
        _write_attr_int(d, "__complex__", 1)
    L.H5Dclose(d), L.H5Sclose(sp)


def _read_dataset(g, key):
    L = lib()
    d = _ok(L.H5Dopen2(g, key.encode(), H5P_DEFAULT), f"open {key}")
    t = L.H5Dget_type(d)
    cls = L.H5Tget_class(t)
    if cls == H5T_STRING:
        size = L.H5Tget_size(t)
        buf = C.create_string_buffer(size + 1)
        L.H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf)
        L.H5Tclose(t), L.H5Dclose(d)
        return buf.value.decode()
    sp = L.H5Dget_space(d)
    nd = L.H5Sget_simple_extent_ndims(sp)
    dims = (hsize_t * max(nd, 1))()
    if nd > 0:
        L.H5Sget_simple_extent_dims(sp, dims, None)
    shape = tuple(int(dims[i]) for i in range(nd))
    dtype, tt = (np.int64, _t("H5T_NATIVE_LLONG_g")) if cls == H5T_INTEGER else (np.float64, _t("H5T_NATIVE_DOUBLE_g"))
    a = np.zeros(shape, dtype=dtype)
    if a.size:
        _ok(L.H5Dread(d, tt, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data_as(C.c_void_p)), f"read {key}")
    cplx = L.H5Aexists(d, b"__complex__") > 0
    L.H5Sclose(sp), L.H5Tclose(t), L.H5Dclose(d)
    if cplx:
        a = a[..., 0] + 1j * a[..., 1]
    if a.ndim == 0:
        return a.item()
    return a


def _write(g, key, value):
    L = lib()
    if value is None or isinstance(value, (dict, list, tuple)) or hasattr(value, "__reduce_to_dict__"):
        sub = _ok(L.H5Gcreate2(g, key.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"group {key}")
        if value is None:
            _write_attr_str(sub, "Format", "None")
        elif isinstance(value, dict):
            for k, v in value.items():
                _write(sub, str(k), v)
        elif isinstance(value, (list, tuple)):
            _write_attr_str(sub, "Format", "list" if isinstance(value, list) else "tuple")
            for i, v in enumerate(value):
                _write(sub, str(i), v)
        else:
            _write_attr_str(sub, "Format", type(value).__name__)
            for k, v in value.__reduce_to_dict__().items():
                _write(sub, str(k), v)
        L.H5Gclose(sub)
    else:
        _write_dataset(g, key, value)


def _read(g, key):
    L = lib()
    o = _ok(L.H5Oopen(g, key.encode(), H5P_DEFAULT), f"open {key}")
    kind = L.H5Iget_type(o)
    L.H5Oclose(o)
    if kind == H5I_DATASET:
        return _read_dataset(g, key)
    sub = _ok(L.H5Gopen2(g, key.encode(), H5P_DEFAULT), f"group {key}")
    fmt = _read_attr_str(sub, "Format")
    names = _link_names(sub)
    if fmt == "None":
        out = None
    elif fmt in ("list", "tuple"):
        items = [_read(sub, str(i)) for i in range(len(names))]
        out = items if fmt == "list" else tuple(items)
    else:
        d = {k: _read(sub, k) for k in names}
        if fmt is None:
            out = d
        elif fmt in _registry:
            out = _registry[fmt].__factory_from_dict__(fmt, d)
        else:
            raise H5Error(f"HDFArchive: no class registered for Format {fmt!r}")
    L.H5Gclose(sub)
    return out


class HDFArchive:
    """A[key] = value writes, A[key] reads; mode "w" creates (truncating), "r" reads."""

    def __init__(self, path, mode="r"):
        L = lib()
        self.path = str(path)
        if mode == "w":
            self._f = _ok(L.H5Fcreate(self.path.encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), f"create {path}")
        elif mode == "r":
            self._f = _ok(L.H5Fopen(self.path.encode(), H5F_ACC_RDONLY, H5P_DEFAULT), f"open {path}")
        else:
            raise ValueError(f"HDFArchive: mode must be 'r' or 'w', got {mode!r}")

    def __setitem__(self, key, value):
        _write(self._f, key, value)

    def __getitem__(self, key):
        if lib().H5Lexists(self._f, key.encode(), H5P_DEFAULT) <= 0:
            raise KeyError(key)
        return _read(self._f, key)

    def keys(self):
        return _link_names(self._f)

    def close(self):
        if getattr(self, "_f", None) is not None:
            lib().H5Fclose(self._f)
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()
