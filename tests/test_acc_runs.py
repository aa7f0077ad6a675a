"""The seed-merge tool (acc_runs) on run files written here: grouping by
name tags, the merged sums, the cross-seed energy variance and the SSF
variance terms, checked against the same statistics computed directly from
the per-sample values the files were built from."""
import subprocess

import numpy as np
import pytest

from h5write import write_file

L, NK = 2, 8
T_LIST = np.array([1.0, 0.5])
LABELS = ["xx", "yz"]


def _tool(pyro):
    exe = pyro.host_tool("acc_runs")
    if not exe.exists():
        pytest.skip("acc_runs not built (needs the HDF5 headers at build time)")
    return str(exe)


def _run_file(path, e, samples):
    """One anneal-like run: energy e[T] (one sample per T) and SSF samples
    samples[n_s][n_corr][n_T][n_k][2]."""
    samples = np.asarray(samples, float)
    S = samples.sum(0)
    S2 = (samples ** 2).sum(0)
    n_s = samples.shape[0]
    write_file(path, {
        "/energy/T_list": T_LIST, "/energy/E": np.asarray(e, float), "/energy/E2": np.asarray(e, float) ** 2,
        "/energy/n_samples": np.ones(len(T_LIST), np.uint64),
        "/geometry/index_cell": 8 * np.eye(3, dtype=np.int64),
        "/geometry/recip_vectors": np.pi / 4 * np.eye(3),
        "/ssf/T_list": np.sort(T_LIST), "/ssf/static_corr": S, "/ssf/static_corr_2": S2,
        "/ssf/n_samples": np.full(len(T_LIST), n_s, np.uint64),
        "/ssf/corr_lookup": LABELS, "/ssf/sl_positions": np.arange(48, dtype=np.int64).reshape(16, 3),
    }, {"/ssf": {"n_spins": 16 * L ** 3, "k_dims": [L, L, L]}})


def test_merge_against_per_sample_statistics(pyro, tmp_path):
    from pyro_cmc_amd.h5lite import H5File

    rng = np.random.default_rng(0)
    K, n_s = 3, 5
    energies = rng.normal(size=(K, len(T_LIST)))
    samples = rng.normal(size=(K, n_s, len(LABELS), len(T_LIST), NK, 2))
    names = ["run_L=2_J1=1_J2=0_B=0,0,0_seed=1_Tc=0.5_sw=4_", "run_L=2_J1=1_J2=0.0_B=0,0,0_seed=a2_Tc=0.5_sw=4_",
             "run_L=2_J1=1_J2=0_B=0,0,0_seed=3f_Tc=0.50_sw=4_"]
    for s, name in enumerate(names):
        _run_file(tmp_path / f"{name}.out.h5", energies[s], samples[s])
    # another parameter set, a file without seed, and a stale merge
    _run_file(tmp_path / "run_L=2_J1=1_J2=0.1_B=0,0,0_seed=9_Tc=0.5_sw=4_.out.h5", energies[0], samples[0])
    _run_file(tmp_path / "run_L=2_J1=1_J2=0_nosd_.out.h5", energies[0], samples[0])
    (tmp_path / "old_merge3_.avg.h5").write_bytes(b"stale")

    r = subprocess.run([_tool(pyro), str(tmp_path), "--delete-avg"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "no seed= tag" in r.stderr and not (tmp_path / "old_merge3_.avg.h5").exists()
    assert "in 2 group(s)" in r.stdout
    # named after the group's first file in sorted order, as the reference does
    merged = tmp_path / "run_L=2_J1=1_J2=0.0_B=0,0,0_merge3_Tc=0.5_sw=4_.avg.h5"
    single = tmp_path / "run_L=2_J1=1_J2=0.1_B=0,0,0_merge1_Tc=0.5_sw=4_.avg.h5"
    assert merged.exists() and single.exists()

    with H5File(merged) as f:
        np.testing.assert_allclose(f.read("/energy/E"), energies.sum(0), rtol=1e-14)
        np.testing.assert_array_equal(f.read("/energy/n_samples"), [K, K])
        assert int(f.read("/energy/n_seeds")) == K
        np.testing.assert_allclose(f.read("/energy/e_mean"), energies.mean(0), rtol=1e-13)
        np.testing.assert_allclose(f.read("/energy/var"), energies.var(0), rtol=1e-10)
        np.testing.assert_allclose(f.read("/energy/C"), energies.var(0) / T_LIST ** 2, rtol=1e-10)
        per_seed_mean = samples.mean(1)                          # [K][corr][T][k][2]
        np.testing.assert_allclose(f.read("/ssf/static_corr"), samples.sum((0, 1)), rtol=1e-13)
        np.testing.assert_allclose(f.read("/ssf/static_corr_2"), (samples.sum(1) ** 2).sum(0), rtol=1e-13)
        np.testing.assert_array_equal(f.read("/ssf/n_samples"), [K * n_s] * 2)
        np.testing.assert_allclose(f.read("/ssf/var_inter"), per_seed_mean.var(0), rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(f.read("/ssf/S_mean"), samples.sum((0, 1))[..., 0] / (K * n_s), rtol=1e-12)
        np.testing.assert_allclose(f.read("/ssf/var_intra"), samples.var(1).sum(0), rtol=1e-9, atol=1e-12)
        assert f.read("/ssf/corr_lookup") == LABELS
        np.testing.assert_array_equal(f.attr("/ssf", "k_dims"), [L, L, L])
        assert int(f.attr("/ssf", "n_spins")) == 16 * L ** 3
        np.testing.assert_array_equal(f.read("/geometry/index_cell"), 8 * np.eye(3))
    with H5File(single) as f:
        np.testing.assert_allclose(f.read("/energy/var"), np.zeros(2), atol=1e-15)
        np.testing.assert_allclose(f.read("/ssf/var_inter"), 0.0, atol=1e-12)


def test_mismatched_temperatures_fail(pyro, tmp_path):
    rng = np.random.default_rng(1)
    s = rng.normal(size=(2, len(LABELS), 2, NK, 2))
    _run_file(tmp_path / "a_L=2_seed=1_.out.h5", [1.0, 2.0], s)
    global T_LIST
    saved = T_LIST
    try:
        T_LIST = np.array([1.0, 0.25])
        _run_file(tmp_path / "a_L=2_seed=2_.out.h5", [1.0, 2.0], s)
    finally:
        T_LIST = saved
    r = subprocess.run([_tool(pyro), str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "T_list mismatch" in r.stderr
