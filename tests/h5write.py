"""Small HDF5 writer for tests (ctypes on the same libhdf5 as
pyro_cmc_amd.h5lite): numeric datasets, UTF-8 string datasets and integer
attributes, enough to build run files without a GPU."""
import ctypes

import numpy as np

from pyro_cmc_amd import h5lite

_TYPES = {np.dtype(np.float64): "H5T_NATIVE_DOUBLE_g", np.dtype(np.int64): "H5T_NATIVE_INT64_g",
          np.dtype(np.uint64): "H5T_NATIVE_UINT64_g", np.dtype(np.int32): "H5T_NATIVE_INT32_g"}


def _lib():
    lib = h5lite._load()
    hid = h5lite.hid_t
    for name, res, args in [
        ("H5Fcreate", hid, [ctypes.c_char_p, ctypes.c_uint, hid, hid]),
        ("H5Gcreate2", hid, [hid, ctypes.c_char_p, hid, hid, hid]),
        ("H5Screate_simple", hid, [ctypes.c_int, ctypes.POINTER(h5lite.hsize_t), ctypes.c_void_p]),
        ("H5Screate", hid, [ctypes.c_int]),
        ("H5Dcreate2", hid, [hid, ctypes.c_char_p, hid, hid, hid, hid, hid]),
        ("H5Dwrite", ctypes.c_int, [hid, hid, hid, hid, hid, ctypes.c_void_p]),
        ("H5Acreate2", hid, [hid, ctypes.c_char_p, hid, hid, hid, hid]),
        ("H5Awrite", ctypes.c_int, [hid, hid, ctypes.c_void_p]),
    ]:
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    return lib


def write_file(path, datasets, attrs=None):
    """datasets: {"/group/name": ndarray or list[str]}; attrs:
    {"/group": {"name": int or list[int]}} (a plain int is a scalar)."""
    lib = _lib()
    g = lambda n: h5lite.hid_t.in_dll(lib, n).value  # noqa: E731
    f = lib.H5Fcreate(str(path).encode(), 2, 0, 0)
    groups = {}

    def group(name):
        if name not in groups:
            groups[name] = lib.H5Gcreate2(f, name.encode(), 0, 0, 0)
        return groups[name]

    for full, data in datasets.items():
        gname, dname = full.rsplit("/", 1)
        loc = group(gname)
        if isinstance(data, list) and data and isinstance(data[0], str):
            st = lib.H5Tcopy(g("H5T_C_S1_g"))
            lib.H5Tset_size(st, h5lite.H5T_VARIABLE)
            lib.H5Tset_cset(st, 1)
            arr = (ctypes.c_char_p * len(data))(*[s.encode() for s in data])
            dims = (h5lite.hsize_t * 1)(len(data))
            sp = lib.H5Screate_simple(1, dims, None)
            ds = lib.H5Dcreate2(loc, dname.encode(), st, sp, 0, 0, 0)
            lib.H5Dwrite(ds, st, 0, 0, 0, arr)
            lib.H5Tclose(st)
        else:
            a = np.ascontiguousarray(data)
            t = g(_TYPES[a.dtype])
            if a.ndim == 0:
                sp = lib.H5Screate(0)
            else:
                dims = (h5lite.hsize_t * a.ndim)(*a.shape)
                sp = lib.H5Screate_simple(a.ndim, dims, None)
            ds = lib.H5Dcreate2(loc, dname.encode(), t, sp, 0, 0, 0)
            lib.H5Dwrite(ds, t, 0, 0, 0, a.ctypes.data_as(ctypes.c_void_p))
        lib.H5Dclose(ds)
        lib.H5Sclose(sp)
    for gname, items in (attrs or {}).items():
        loc = group(gname)
        for name, v in items.items():
            scalar = np.isscalar(v)
            a = np.ascontiguousarray(np.atleast_1d(v), np.int32)
            if scalar:
                sp = lib.H5Screate(0)
            else:
                dims = (h5lite.hsize_t * 1)(len(a))
                sp = lib.H5Screate_simple(1, dims, None)
            at = lib.H5Acreate2(loc, name.encode(), g("H5T_NATIVE_INT_g"), sp, 0, 0)
            lib.H5Awrite(at, g("H5T_NATIVE_INT_g"), a.ctypes.data_as(ctypes.c_void_p))
            lib.H5Aclose(at)
            lib.H5Sclose(sp)
    for gid in groups.values():
        lib.H5Gclose(gid)
    lib.H5Fclose(f)
