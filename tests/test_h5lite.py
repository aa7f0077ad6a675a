"""The ctypes HDF5 reader against files written here through the same
libhdf5 (numeric datasets of each stored type, UTF-8 variable-length
strings, scalar and 1-D attributes, group listings)."""
import numpy as np
import pytest


def test_round_trip(tmp_path):
    try:
        from h5write import write_file
        from pyro_cmc_amd.h5lite import H5File
    except OSError as e:  # pragma: no cover - libhdf5 missing
        pytest.skip(f"libhdf5 unavailable: {e}")
    data = {
        "d": np.arange(24, dtype=np.float64).reshape(2, 3, 4) / 7,
        "i": np.array([[1, -2, 3], [4, 5, -6]], np.int64),
        "u": np.array([7, 2 ** 40], np.uint64),
        "p": np.array([[1, 1, 1], [-1, 3, 5]], np.int32),
    }
    labels = ["xx", "yz", "μν"]
    path = tmp_path / "t.h5"
    write_file(path, {**{f"/grp/{k}": v for k, v in data.items()}, "/grp/labels": labels},
               {"/grp": {"n_spins": 1024, "k_dims": [4, 5, 6]}})
    with H5File(path) as f:
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
