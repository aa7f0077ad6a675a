"""Luttinger-Tisza spectrum (luttinger_tisza.cpp:31-181).

The oracle's restatement is pinned on the CPU by closed forms and a brute
force that shares none of its displacement tables:
- J1-only Heisenberg: the pyrochlore flat band, lambda_min(k) = -2 J1 at
  every k (E/spin = -J1, the Moessner-Chalker ground-state energy);
- k = 0, Delta = 1: the uniform state is an eigenvector with eigenvalue
  6 J1 + 12 J2 + 12 J3 (coordination numbers 6, 12, 6 + 6);
- Bloch's theorem: the union of the spectra of M(k) over the L^3 grid is the
  spectrum of the real-space coupling matrix of the L^3 supercell, built
  here from a distance search of the positions (with XXZ anisotropy);
- the LT bound: every spin configuration has E/spin >= lambda_min / 2.
The GPU tests compare pyro_lt_spectrum with the oracle (1e-12 relative on
every lambda_min(k); the eigenvector through ||M v - lambda v||) and the
ltgs driver's HDF5 layout with the reference's (luttinger_tisza.cpp:209-304).
"""
import subprocess

import numpy as np
import pytest

from helpers import brute_coupling_matrix, brute_energy_per_cell, random_spins

CASES = [(1.0, 0.0, 0.0, 1.0), (1.0, -0.25, 0.15, 1.0), (1.0, 0.1, -0.2, 0.35), (-0.7, 0.3, 0.05, 2.5)]


def test_oracle_flat_band(oracle):
    lam, k_star, v = oracle.lt_spectrum(4, 1.3)
    np.testing.assert_allclose(lam, -2 * 1.3, atol=1e-12)
    assert abs(np.linalg.norm(v) - 1) < 1e-12


def test_oracle_uniform_state_at_gamma(oracle):
    J1, J2, J3 = 0.8, -0.3, 0.45
    M = oracle.lt_matrices(np.zeros((1, 3)), oracle.lt_bonds(J1, J2, J3))[0]
    for a in range(3):
        u = np.zeros(48)
        u[a::3] = 1.0
        np.testing.assert_allclose(M @ u, (6 * J1 + 12 * J2 + 12 * J3) * u, atol=1e-13)
    np.testing.assert_allclose(M, M.conj().T, atol=1e-15)


@pytest.mark.parametrize("J1,J2,J3,Delta", CASES)
def test_oracle_bloch_union_equals_real_space_spectrum(oracle, J1, J2, J3, Delta):
    L = 2
    K, Ls = oracle.lt_k_grid(L)
    M = oracle.lt_matrices(oracle.lt_q(K, Ls), oracle.lt_bonds(J1, J2, J3, Delta))
    union = np.sort(np.linalg.eigvalsh(M).ravel())
    H = brute_coupling_matrix(L, J1, J2, J3, Delta)
    np.testing.assert_allclose(union, np.linalg.eigvalsh(H), atol=1e-11)


def test_lt_bound_holds_for_random_states(oracle):
    L, (J1, J2, J3, Delta) = 2, CASES[2]
    lam, _, _ = oracle.lt_spectrum(L, J1, J2, J3, Delta)
    rng = np.random.default_rng(5)
    for _ in range(20):
        S = random_spins(16 * L ** 3, rng)
        e_spin = brute_energy_per_cell(S, L, J1, J2, J3, Delta) / 16
        assert e_spin >= lam.min() / 2 - 1e-12


def _check_against_oracle(gpu, oracle, L, J1, J2, J3, Delta):
    lam, k_star, v = gpu.lt_spectrum(L, J1, J2, J3, Delta)
    lam_o, _, _ = oracle.lt_spectrum(L, J1, J2, J3, Delta)
    scale = abs(J1) + abs(J2) + abs(J3)
    np.testing.assert_allclose(lam, lam_o, rtol=0, atol=1e-12 * 30 * scale * max(1.0, abs(Delta)))
    # k*: a k of the minimum (ties between symmetric k differ in the last bits)
    Ls = np.array([L] * 3 if np.isscalar(L) else L)
    flat = (k_star[0] * Ls[1] + k_star[1]) * Ls[2] + k_star[2]
    assert lam[flat] <= lam.min() + 1e-12 * scale
    M = oracle.lt_matrices(oracle.lt_q(k_star, Ls)[None], oracle.lt_bonds(J1, J2, J3, Delta))[0]
    assert abs(np.linalg.norm(v) - 1) < 1e-12
    assert np.linalg.norm(M @ v - lam[flat] * v) < 1e-11 * scale * max(1.0, abs(Delta))


@pytest.mark.gpu
@pytest.mark.parametrize("L", [2, 3, 5])
@pytest.mark.parametrize("J1,J2,J3,Delta", CASES)
def test_lt_spectrum_matches_oracle(gpu, oracle, L, J1, J2, J3, Delta):
    _check_against_oracle(gpu, oracle, L, J1, J2, J3, Delta)


@pytest.mark.gpu
def test_lt_spectrum_rectangular_grid(gpu, oracle):
    _check_against_oracle(gpu, oracle, (2, 3, 4), *CASES[1])


@pytest.mark.gpu
def test_lt_flat_band_full_size(gpu):
    """L = 24 (13,824 k, three rocSOLVER batches' worth of matrices)."""
    lam, k_star, v = gpu.lt_spectrum(24, 1.0)
    np.testing.assert_allclose(lam, -2.0, atol=1e-11)


@pytest.mark.gpu
def test_lt_rejects_bad_arguments(gpu):
    with pytest.raises(gpu.PyroError):
        gpu.lt_spectrum((0, 2, 2), 1.0)
    with pytest.raises(gpu.PyroError):
        gpu.lt_spectrum(2, float("nan"))


@pytest.mark.gpu
def test_ltgs_file_layout(gpu, oracle, tmp_path):
    from pyro_cmc_amd.h5lite import H5File

    exe = gpu.host_tool("ltgs")
    if not exe.exists():
        pytest.skip("ltgs not built (needs the HDF5 headers at build time)")
    L, J1, J2, J3, Delta = 4, 1.0, -0.25, 0.15, 0.6
    r = subprocess.run([str(exe), "-o", str(tmp_path), str(L), "--J1", str(J1), "--J2", str(J2), "--J3", str(J3),
                        "--Delta", str(Delta)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "LT minimum" in r.stdout and "k*  idx" in r.stdout
    files = list(tmp_path.glob("*ltgs.out.h5"))
    assert [f.name for f in files] == ["L=4_J1=1_J2=-0.25_J3=0.15_Delta=0.6_ltgs.out.h5"]
    lam_o, _, _ = oracle.lt_spectrum(L, J1, J2, J3, Delta)
    with H5File(files[0]) as f:
        lam = f.read("eigenvalue_map")
        np.testing.assert_allclose(lam, lam_o, atol=1e-11)
        ks = f.read("k_star_idx")
        assert ks.dtype == np.int64 and ks.shape == (3,)
        flat = (ks[0] * L + ks[1]) * L + ks[2]
        assert float(f.read("E_min")) == pytest.approx(lam[flat] / 2, abs=1e-15)
        assert float(f.read("E_min")) == pytest.approx(lam_o.min() / 2, abs=1e-11)
        np.testing.assert_allclose(f.read("k_star_vec"), oracle.lt_q(ks, np.array([L] * 3)), atol=1e-15)
        ev = f.read("eigenvector")
        assert ev.shape == (48, 2)
        v = ev[:, 0] + 1j * ev[:, 1]
        M = oracle.lt_matrices(oracle.lt_q(ks, np.array([L] * 3))[None], oracle.lt_bonds(J1, J2, J3, Delta))[0]
        assert np.linalg.norm(M @ v - lam[flat] * v) < 1e-10
        np.testing.assert_array_equal(f.read("geometry/index_cell"), 8 * np.eye(3))
        np.testing.assert_array_equal(f.attr("/", "k_dims"), [L, L, L])


def test_ltgs_prints_without_output_dir_flag_parse(pyro):
    exe = pyro.host_tool("ltgs")
    if not exe.exists():
        pytest.skip("ltgs not built")
    r = subprocess.run([str(exe), "4"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "--J1" in r.stderr
