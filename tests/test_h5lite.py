"""The ctypes HDF5 reader against files written here through the same
libhdf5 (numeric datasets of each stored type, UTF-8 variable-length
strings, scalar and 1-D attributes, group listings)."""
import ctypes

import numpy as np
import pytest


@pytest.fixture(scope="module")
def h5():
    try:
        from pyro_cmc_amd import h5lite

        lib = h5lite._load()
    except OSError as e:
        pytest.skip(f"libhdf5 unavailable: {e}")
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
    return h5lite, lib


def _write(h5, path):
    h5lite, lib = h5
    g = lambda n: h5lite.hid_t.in_dll(lib, n).value  # noqa: E731
    f = lib.H5Fcreate(str(path).encode(), 2, 0, 0)  # H5F_ACC_TRUNC
    grp = lib.H5Gcreate2(f, b"/grp", 0, 0, 0)
    data = {
        "d": (np.arange(24, dtype=np.float64).reshape(2, 3, 4) / 7, "H5T_NATIVE_DOUBLE_g"),
        "i": (np.array([[1, -2, 3], [4, 5, -6]], np.int64), "H5T_NATIVE_INT64_g"),
        "u": (np.array([7, 2 ** 40], np.uint64), "H5T_NATIVE_UINT64_g"),
        "p": (np.array([[1, 1, 1], [-1, 3, 5]], np.int32), "H5T_NATIVE_INT32_g"),
    }
    for name, (a, t) in data.items():
        dims = (h5lite.hsize_t * a.ndim)(*a.shape)
        sp = lib.H5Screate_simple(a.ndim, dims, None)
        ds = lib.H5Dcreate2(grp, name.encode(), g(t), sp, 0, 0, 0)
        assert lib.H5Dwrite(ds, g(t), 0, 0, 0, a.ctypes.data_as(ctypes.c_void_p)) >= 0
        lib.H5Dclose(ds)
        lib.H5Sclose(sp)
    st = lib.H5Tcopy(g("H5T_C_S1_g"))
    lib.H5Tset_size(st, h5lite.H5T_VARIABLE)
    lib.H5Tset_cset(st, 1)  # UTF-8
    labels = ["xx", "yz", "μν"]
    arr = (ctypes.c_char_p * 3)(*[s.encode() for s in labels])
    dims = (h5lite.hsize_t * 1)(3)
    sp = lib.H5Screate_simple(1, dims, None)
    ds = lib.H5Dcreate2(grp, b"labels", st, sp, 0, 0, 0)
    assert lib.H5Dwrite(ds, st, 0, 0, 0, arr) >= 0
    lib.H5Dclose(ds)
    lib.H5Sclose(sp)
    lib.H5Tclose(st)
    one = ctypes.c_int(1024)
    sp = lib.H5Screate(0)  # H5S_SCALAR
    at = lib.H5Acreate2(grp, b"n_spins", g("H5T_NATIVE_INT_g"), sp, 0, 0)
    lib.H5Awrite(at, g("H5T_NATIVE_INT_g"), ctypes.byref(one))
    lib.H5Aclose(at)
    lib.H5Sclose(sp)
    kd = np.array([4, 5, 6], np.int32)
    sp = lib.H5Screate_simple(1, dims, None)
    at = lib.H5Acreate2(grp, b"k_dims", g("H5T_NATIVE_INT_g"), sp, 0, 0)
    lib.H5Awrite(at, g("H5T_NATIVE_INT_g"), kd.ctypes.data_as(ctypes.c_void_p))
    lib.H5Aclose(at)
    lib.H5Sclose(sp)
    lib.H5Gclose(grp)
    lib.H5Fclose(f)
    return {k: v[0] for k, v in data.items()}, labels


def test_round_trip(h5, tmp_path):
    h5lite, _ = h5
    path = tmp_path / "t.h5"
    data, labels = _write(h5, path)
    with h5lite.H5File(path) as f:
        assert f.keys("/") == ["grp"]
        assert sorted(f.keys("/grp")) == ["d", "i", "labels", "p", "u"]
        assert f.exists("/grp/d") and not f.exists("/grp/zz") and not f.exists("/nope/d")
        for k, a in data.items():
            got = f.read(f"/grp/{k}")
            assert got.dtype == a.dtype and got.shape == a.shape
            np.testing.assert_array_equal(got, a)
        assert f.read("/grp/labels") == labels
        assert int(f.attr("/grp", "n_spins")) == 1024 and f.attr("/grp", "n_spins").shape == ()
        np.testing.assert_array_equal(f.attr("/grp", "k_dims"), [4, 5, 6])
        with pytest.raises(KeyError):
            f.read("/grp/missing")
