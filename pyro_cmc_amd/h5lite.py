"""Minimal read-only HDF5 access through ctypes on the system libhdf5.

Reads the files the anneal/temper drivers write (numeric datasets,
variable-length UTF-8 string datasets, numeric attributes, group listings)
without h5py, which this environment does not provide.

    with H5File(path) as f:
        S = f.read("/ssf/static_corr")          # numpy array
        labels = f.read("/ssf/corr_lookup")     # list[str]
        n_spins = f.attr("/ssf", "n_spins")     # numpy array (shape () for scalars)
        f.keys("/ssf")                          # member names
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_CANDIDATES = [
    os.environ.get("PYRO_HDF5_LIB", ""),
    "/opt/conda/lib/libhdf5.so.103",
    "/opt/conda/lib/libhdf5.so",
    "libhdf5.so",
]

hid_t = ctypes.c_int64
herr_t = ctypes.c_int
hsize_t = ctypes.c_uint64

H5F_ACC_RDONLY = 0
H5P_DEFAULT = 0
H5S_ALL = 0
H5T_VARIABLE = ctypes.c_size_t(-1).value
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3
H5T_SGN_NONE = 0
H5_INDEX_NAME, H5_ITER_INC = 0, 0

_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    err = None
    for cand in _CANDIDATES:
        if not cand:
            continue
        try:
            lib = ctypes.CDLL(cand)
            break
        except OSError as e:
            err = e
    else:
        raise OSError(f"libhdf5 not found ({err})")
    for name, res, args in [
        ("H5open", herr_t, []),
        ("H5Fopen", hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t]),
        ("H5Fclose", herr_t, [hid_t]),
        ("H5Dopen2", hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        ("H5Dclose", herr_t, [hid_t]),
        ("H5Dget_space", hid_t, [hid_t]),
        ("H5Dget_type", hid_t, [hid_t]),
        ("H5Dread", herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        ("H5Dvlen_reclaim", herr_t, [hid_t, hid_t, hid_t, ctypes.c_void_p]),
        ("H5Sget_simple_extent_ndims", ctypes.c_int, [hid_t]),
        ("H5Sget_simple_extent_dims", ctypes.c_int, [hid_t, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]),
        ("H5Sclose", herr_t, [hid_t]),
        ("H5Tget_class", ctypes.c_int, [hid_t]),
        ("H5Tget_size", ctypes.c_size_t, [hid_t]),
        ("H5Tget_sign", ctypes.c_int, [hid_t]),
        ("H5Tis_variable_str", ctypes.c_int, [hid_t]),
        ("H5Tcopy", hid_t, [hid_t]),
        ("H5Tset_size", herr_t, [hid_t, ctypes.c_size_t]),
        ("H5Tget_cset", ctypes.c_int, [hid_t]),
        ("H5Tset_cset", herr_t, [hid_t, ctypes.c_int]),
        ("H5Tclose", herr_t, [hid_t]),
        ("H5Aopen_by_name", hid_t, [hid_t, ctypes.c_char_p, ctypes.c_char_p, hid_t, hid_t]),
        ("H5Aget_space", hid_t, [hid_t]),
        ("H5Aget_type", hid_t, [hid_t]),
        ("H5Aread", herr_t, [hid_t, hid_t, ctypes.c_void_p]),
        ("H5Aclose", herr_t, [hid_t]),
        ("H5Lexists", ctypes.c_int, [hid_t, ctypes.c_char_p, hid_t]),
        ("H5Gopen2", hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        ("H5Gclose", herr_t, [hid_t]),
        ("H5Lget_name_by_idx", ctypes.c_ssize_t,
         [hid_t, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, hsize_t, ctypes.c_char_p, ctypes.c_size_t, hid_t]),
    ]:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.H5open() < 0:
        raise OSError("H5open failed")
    # errors come back as return codes; keep HDF5's own stack printout quiet
    lib.H5Eset_auto2.argtypes = [hid_t, ctypes.c_void_p, ctypes.c_void_p]
    lib.H5Eset_auto2(0, None, None)
    _lib = lib
    return lib


def _native(lib, cls: int, size: int, signed: bool):
    """(numpy dtype, native HDF5 memory type id) for a stored numeric type."""
    g = lambda n: hid_t.in_dll(lib, n).value  # noqa: E731
    if cls == H5T_FLOAT:
        return (np.float64, g("H5T_NATIVE_DOUBLE_g")) if size == 8 else (np.float32, g("H5T_NATIVE_FLOAT_g"))
    if cls == H5T_INTEGER:
        table = {
            (8, True): (np.int64, "H5T_NATIVE_INT64_g"),
            (8, False): (np.uint64, "H5T_NATIVE_UINT64_g"),
            (4, True): (np.int32, "H5T_NATIVE_INT32_g"),
            (4, False): (np.uint32, "H5T_NATIVE_UINT_g"),
            (1, True): (np.int8, "H5T_NATIVE_INT8_g"),
            (1, False): (np.uint8, "H5T_NATIVE_UINT8_g"),
        }
        dt, name = table[(size, signed)]
        return dt, g(name)
    raise TypeError(f"unsupported HDF5 type class {cls}")


def _shape(lib, space: int) -> tuple[int, ...]:
    nd = lib.H5Sget_simple_extent_ndims(space)
    if nd <= 0:
        return ()
    dims = (hsize_t * nd)()
    lib.H5Sget_simple_extent_dims(space, dims, None)
    return tuple(int(d) for d in dims)


class H5File:
    def __init__(self, path: str | os.PathLike):
        self.lib = _load()
        self.path = Path(path)
        self.fid = self.lib.H5Fopen(str(self.path).encode(), H5F_ACC_RDONLY, H5P_DEFAULT)
        if self.fid < 0:
            raise OSError(f"cannot open {self.path}")

    def close(self):
        if self.fid >= 0:
            self.lib.H5Fclose(self.fid)
            self.fid = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def exists(self, name: str) -> bool:
        parts = [p for p in name.split("/") if p]
        cur = ""
        for p in parts:
            cur += "/" + p
            if self.lib.H5Lexists(self.fid, cur.encode(), H5P_DEFAULT) <= 0:
                return False
        return True

    def keys(self, group: str = "/") -> list[str]:
        out = []
        buf = ctypes.create_string_buffer(1024)
        i = 0
        while True:
            n = self.lib.H5Lget_name_by_idx(self.fid, group.encode(), H5_INDEX_NAME, H5_ITER_INC, i, buf, 1024,
                                            H5P_DEFAULT)
            if n < 0:
                break
            out.append(buf.value.decode())
            i += 1
        return out

    def read(self, name: str):
        lib = self.lib
        ds = lib.H5Dopen2(self.fid, name.encode(), H5P_DEFAULT)
        if ds < 0:
            raise KeyError(name)
        space = lib.H5Dget_space(ds)
        ftype = lib.H5Dget_type(ds)
        try:
            shape = _shape(lib, space)
            cls = lib.H5Tget_class(ftype)
            if cls == H5T_STRING:
                if not lib.H5Tis_variable_str(ftype):
                    raise TypeError("only variable-length strings are supported")
                n = int(np.prod(shape)) if shape else 1
                mt = lib.H5Tcopy(hid_t.in_dll(lib, "H5T_C_S1_g").value)
                lib.H5Tset_size(mt, H5T_VARIABLE)
                lib.H5Tset_cset(mt, lib.H5Tget_cset(ftype))  # no ASCII <-> UTF-8 conversion exists
                arr = (ctypes.c_char_p * n)()
                if lib.H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr) < 0:
                    raise OSError(f"H5Dread {name}")
                out = [s.decode() for s in arr]
                lib.H5Dvlen_reclaim(mt, space, H5P_DEFAULT, arr)
                lib.H5Tclose(mt)
                return out
            dt, mt = _native(lib, cls, lib.H5Tget_size(ftype), lib.H5Tget_sign(ftype) != H5T_SGN_NONE)
            out = np.empty(shape, dtype=dt)
            if out.size and lib.H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                        out.ctypes.data_as(ctypes.c_void_p)) < 0:
                raise OSError(f"H5Dread {name}")
            return out
        finally:
            lib.H5Tclose(ftype)
            lib.H5Sclose(space)
            lib.H5Dclose(ds)

    def attr(self, obj: str, name: str) -> np.ndarray:
        lib = self.lib
        a = lib.H5Aopen_by_name(self.fid, obj.encode(), name.encode(), H5P_DEFAULT, H5P_DEFAULT)
        if a < 0:
            raise KeyError(f"{obj}@{name}")
        space = lib.H5Aget_space(a)
        ftype = lib.H5Aget_type(a)
        try:
            dt, mt = _native(lib, lib.H5Tget_class(ftype), lib.H5Tget_size(ftype),
                             lib.H5Tget_sign(ftype) != H5T_SGN_NONE)
            out = np.empty(_shape(lib, space), dtype=dt)
            if lib.H5Aread(a, mt, out.ctypes.data_as(ctypes.c_void_p)) < 0:
                raise OSError(f"H5Aread {obj}@{name}")
            return out
        finally:
            lib.H5Tclose(ftype)
            lib.H5Sclose(space)
            lib.H5Aclose(a)
