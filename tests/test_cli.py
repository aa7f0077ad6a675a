"""Host drivers (anneal, temper, spiral_byhand) and their HDF5 output.

CPU tests cover argument handling and the seed derivation.  GPU tests run
each driver on a small lattice and check:
  * the HDF5 layout of anneal.cpp:229-245 / ssf_manager.hpp:413-641 /
    MC.cpp:416-454 (datasets, shapes, dtypes, attributes, T order);
  * that the file holds exactly what the same sequence of library calls
    gives (a replay through the Python bindings), so the drivers add no
    arithmetic of their own;
  * spiral_byhand's closed forms (S^yy = S^zz = N^2/4 at +-Q, S^yz = -+iN^2/4,
    S^xx = 0) and brute-force pair energies of the saved spins;
  * the saved spin DFT against the oracle's direct DFT.
"""
import re
import subprocess

import numpy as np
import pytest

from helpers import brute_energy_per_cell, pyro_positions

TOOLS = ("anneal", "temper", "spiral_byhand", "acc_runs", "fieldcool", "ltgs")


def _tool(pyro, name):
    exe = pyro.host_tool(name)
    if not exe.exists():
        pytest.skip(f"{name} not built (needs the HDF5 headers at build time)")
    return str(exe)


def _run(cmd, timeout=120):
    print("RUN", " ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)


# ---------------------------------------------------------------------------
# CPU
# ---------------------------------------------------------------------------
def test_seed_key_is_splitmix64(pyro):
    # SplitMix64 published first outputs for states 0 and 1234567
    assert pyro.seed_key(0) == 0xE220A8397B1DCDAF
    assert pyro.seed_key(1234567) == 0x599ED017FB08FC85


@pytest.mark.parametrize("name", TOOLS)
def test_missing_required_arguments(pyro, name):
    r = _run([_tool(pyro, name)])
    assert r.returncode == 1
    assert "missing required argument" in r.stderr
    assert "usage:" in r.stderr


@pytest.mark.parametrize("name", ["bench_lifted", "bench_speed"])
def test_harness_rejects_unknown_option(pyro, name):
    """The reference's test harnesses exit with status 2 on a bad command
    line (bench_lifted.cpp:172-176, bench_speed.cpp:128-133)."""
    r = _run([_tool(pyro, name), "--no_such_option", "1"])
    assert r.returncode == 2
    assert "unknown option --no_such_option" in r.stderr


def test_unknown_option_and_bad_outdir(pyro, tmp_path):
    exe = _tool(pyro, "anneal")
    r = _run([exe, "-o", str(tmp_path), "-s", "0", "--T_sample", "1", "--J1", "1", "4", "--bogus", "1"])
    assert r.returncode == 1 and "unknown option --bogus" in r.stderr
    r = _run([exe, "-o", str(tmp_path / "nope"), "-s", "0", "--T_sample", "1", "--J1", "1", "4"])
    assert r.returncode == 1 and "Cannot open outdir" in r.stderr


def test_temper_rejects_short_ladder(pyro, tmp_path):
    exe = _tool(pyro, "temper")
    r = _run([exe, "-o", str(tmp_path), "-s", "0", "--T_sample", "1", "0.5", "-R", "2", "--J1", "1", "2"])
    assert r.returncode == 1 and "must exceed T_sample count" in r.stderr
    r = _run([exe, "-o", str(tmp_path), "-s", "0", "--T_sample", "1", "--J1", "1", "--J3", "0.1", "--Q", "0.5", "2"])
    assert r.returncode == 1 and "mutually exclusive" in r.stderr


def test_gpu_test_arguments_parse(pyro, tmp_path):
    """The argument lists of the GPU tests below parse; without a device
    the drivers then stop at the first device operation."""
    if pyro.device_count() > 0:
        pytest.skip("a device is present: the GPU tests run these")
    cases = [
        ("spiral_byhand", ["-o", str(tmp_path), "4", "0", "0", "1", "--J2", "0.3", "--J3", "-0.2"]),
        ("anneal", _anneal_args(tmp_path, True)),
        ("temper", _temper_args(tmp_path, 6, 2, 5, True)),
    ]
    for name, args in cases:
        r = _run([_tool(pyro, name), *args])
        assert r.returncode == 1
        assert "no HIP device available" in r.stderr, r.stderr


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
def _h5(path):
    from pyro_cmc_amd.h5lite import H5File

    return H5File(path)


def _check_ssf_metadata(f, L, labels, T_list):
    assert f.read("/ssf/corr_lookup") == list(labels)
    np.testing.assert_array_equal(f.read("/ssf/T_list"), np.sort(T_list))
    assert f.read("/ssf/n_samples").dtype == np.uint64
    pos = f.read("/ssf/sl_positions")
    assert pos.dtype == np.int64 and pos.shape == (16, 3)
    assert int(f.attr("/ssf", "n_spins")) == 16 * L ** 3
    np.testing.assert_array_equal(f.attr("/ssf", "k_dims"), [L, L, L])
    np.testing.assert_array_equal(f.read("/geometry/index_cell"), 8 * np.eye(3, dtype=np.int64))
    np.testing.assert_allclose(f.read("/geometry/recip_vectors"), 2 * np.pi / 8 * np.eye(3), rtol=0, atol=1e-15)


@pytest.mark.gpu
def test_spiral_byhand_closed_forms(gpu, oracle, tmp_path):
    L, Q = 4, (0, 0, 1)
    r = _run([_tool(gpu, "spiral_byhand"), "-o", str(tmp_path), str(L), *map(str, Q), "--J2", "0.3", "--J3", "-0.2"])
    assert r.returncode == 0, r.stderr
    out = tmp_path / f"L={L}_Q={Q[0]},{Q[1]},{Q[2]}_.out.h5"
    spins = tmp_path / f"L={L}_Q={Q[0]},{Q[1]},{Q[2]}_.spins.h5"
    N = 16 * L ** 3
    with _h5(spins) as f:
        pos = f.read("spin_pos")
        S = f.read("spin_orientation")
    np.testing.assert_array_equal(pos, pyro_positions(L))
    q = 2 * np.pi * np.array(Q) / (8 * L)
    ph = pos @ q
    np.testing.assert_allclose(S, np.stack([0 * ph, np.cos(ph), np.sin(ph)], 1), atol=1e-15)
    with _h5(out) as f:
        _check_ssf_metadata(f, L, ["xx", "yy", "zz"], [0.0])
        Sq = f.read("/ssf/static_corr")
        assert Sq.shape == (3, 1, L ** 3, 2)
        E = f.read("/energy/E")
        np.testing.assert_array_equal(f.read("/energy/n_samples"), [1])
    K = (Q[0] * L + Q[1]) * L + Q[2]
    Km = ((-Q[0]) % L * L + (-Q[1]) % L) * L + (-Q[2]) % L
    expect = np.zeros((3, L ** 3, 2))
    expect[1, K, 0] = expect[1, Km, 0] = N * N / 4
    expect[2, K, 0] = expect[2, Km, 0] = N * N / 4
    np.testing.assert_allclose(Sq[:, 0], expect, rtol=0, atol=1e-8 * N * N)
    e_ref = brute_energy_per_cell(S, L, -1.0, 0.3, -0.2)
    np.testing.assert_allclose(E[0], e_ref, rtol=1e-12, atol=1e-12)


def _anneal_args(tmp_path, lifted=False):
    args = ["-o", str(tmp_path), "-s", "3", "--J1", "1", "--J2", "-0.05", "--J3", "0.05", "--Delta", "1.2",
            "-B", "0", "0.1", "0.2", "--T_sample", "0.5", "0.2", "--T_hot", "3", "--n_steps", "4",
            "-w", "2", "--n_burn_in", "3", "--n_sample", "3", "--save_state", "--save_state_ft", "2"]
    if lifted:
        args.append("--lifted")
    return args


@pytest.mark.gpu
@pytest.mark.parametrize("lifted", [False, True])
def test_anneal_file_matches_library_replay(gpu, oracle, tmp_path, lifted):
    pyro = gpu
    r = _run([_tool(pyro, "anneal"), *_anneal_args(tmp_path, lifted)])
    assert r.returncode == 0, r.stderr
    stem = "run_L=2_J1=1_J2=-0.05_J3=0.05_Delta=1.2_B=0,0.1,0.2_seed=3_Tc=0.2_sw=2_"
    out = tmp_path / (stem + ".out.h5")
    labels = ["xx", "yy", "zz", "xy", "xz", "yz"]
    with _h5(out) as f:
        assert sorted(f.keys("/")) == ["energy", "geometry", "ssf"]
        T_grid = f.read("/energy/T_list")
        E = f.read("/energy/E")
        E2 = f.read("/energy/E2")
        nE = f.read("/energy/n_samples")
        _check_ssf_metadata(f, 2, labels, [0.5, 0.2])
        Sf = f.read("/ssf/static_corr")
        S2f = f.read("/ssf/static_corr_2")
        nS = f.read("/ssf/n_samples")
    assert len(T_grid) == 4 + 2 and np.all(np.diff(T_grid) < 0) and T_grid[0] == 3.0
    assert Sf.shape == S2f.shape == (6, 2, 8, 2)
    np.testing.assert_array_equal(nE, np.ones(6, np.uint64))
    np.testing.assert_array_equal(nS, [3, 3])

    # replay the driver's library calls
    kind = pyro.LIFTED if lifted else pyro.METROPOLIS
    key = pyro.seed_key(3)
    ctx = pyro.Context.pyrochlore(2, 1, pyro.FP64, seed=key)
    ctx.define_J1J2J3(1.0, -0.05, 0.05, 1.2)
    ctx.setup()
    ctx.set_T_ref(1.0)
    ctx.set_field([0, 0.1, 0.2])
    ctx.randomize(key)
    ctx.ssf_configure(labels, 1, True)
    ctx.set_temperatures([3.0])
    ctx.sweep(kind, 1, 3)
    E_rep, S_rep, S2_rep = [], {}, {}
    for T in T_grid:
        ctx.set_temperatures([T])
        ctx.sweep(kind, 1, 2)
        E_rep.append(ctx.energy_per_cell()[0])
        if T in (0.5, 0.2):
            ctx.ssf_reset(0)
            for _ in range(3):
                ctx.sweep(kind, 1, 2)
                ctx.ssf_sample([0])
            S_rep[T], S2_rep[T], n = ctx.ssf_read(0)
            assert n == 3
    np.testing.assert_array_equal(E, E_rep)
    np.testing.assert_array_equal(E2, np.square(E_rep))
    for j, T in enumerate([0.2, 0.5]):
        np.testing.assert_array_equal(Sf[:, j], S_rep[T])
        np.testing.assert_array_equal(S2f[:, j], S2_rep[T])

    # final state dumps
    S_fin = ctx.get_spins(0)
    with _h5(tmp_path / (stem + ".spins.h5")) as f:
        np.testing.assert_array_equal(f.read("spin_orientation"), S_fin)
        np.testing.assert_array_equal(f.read("spin_pos"), pyro_positions(2))
    with _h5(tmp_path / (stem + ".spindft.h5")) as f:
        ft = f.read("spin_ft")
    assert ft.shape == (3, 16, 8, 2)
    orc = oracle.Oracle.pyrochlore(2)
    orc.set_spins(S_fin)
    np.testing.assert_allclose(ft, orc.spin_ft(), rtol=0, atol=1e-12)


def _temper_args(tmp_path, R, L, seed, munu):
    args = ["-o", str(tmp_path), "-s", str(seed), "--J1", "1", "--J2", "-0.05", "--J3", "0.05",
            "--T_sample", "0.4", "0.25", "--T_hot", "2", "-R", str(R), "-w", "2", "--n_burn_in", "4",
            "--n_sample", "7", "--save_state", str(L)]
    if munu:
        args.append("--munu")
    return args


def _oracle_colour_sweeps(o, r, T, colours, n_col, key, p, n_sweeps, n_or=1):
    """n_sweeps counter-mode Metropolis + n_or OR sweeps of oracle o as
    replica r from pass index p; returns the next pass index."""
    for _ in range(n_sweeps):
        o.pass_ctr(0, T, colours, n_col, key, r, p)
        p += 1
        for _ in range(n_or):
            o.pass_ctr(3, T, colours, n_col, key, r, p)
            p += 1
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("munu", [False, True])
def test_temper_file_matches_oracle_replay(gpu, oracle, tmp_path, munu):
    """The whole temper run (temper.cpp:418-508) replayed by the oracle:
    counter-mode colour sweeps of every replica at its rung's temperature,
    the SSF of the T_sample rungs, one energy sample per rung and round,
    and the checkerboard exchange decided as temper.cpp:437-473 decides it.
    Energies, swap table, SSF sums (contracted and sublattice-resolved) and
    the coldest replica's spins of the files equal the replay."""
    pyro = gpu
    R, L, seed, n_rounds = 6, 2, 5, 7
    r = _run([_tool(pyro, "temper"), *_temper_args(tmp_path, R, L, seed, munu)])
    assert r.returncode == 0, r.stderr
    stem = f"run_L={L}_J1=1_J2=-0.05_J3=0.05_Delta=1_B=0,0,0_seed={seed}_Tc=0.25_sw=2_"
    with _h5(tmp_path / (stem + ".out.h5")) as f:
        T_asc = f.read("/energy/T_list")
        E = f.read("/energy/E")
        E2 = f.read("/energy/E2")
        nE = f.read("/energy/n_samples")
        _check_ssf_metadata(f, L, ["xx", "yy", "zz"], [0.4, 0.25])
        Sf = f.read("/ssf/static_corr")
        S2f = f.read("/ssf/static_corr_2")
        nS = f.read("/ssf/n_samples")
        Mf = f.read("/ssf/static_corr_munu") if munu else None
        assert f.exists("/ssf/static_corr_munu") == munu
    assert len(T_asc) == R and np.all(np.diff(T_asc) > 0) and T_asc[-1] == 2.0
    T_grid = T_asc[::-1].copy()  # slot 0 hottest
    assert Sf.shape == (3, 2, L ** 3, 2)
    np.testing.assert_array_equal(nE, np.full(R, n_rounds, np.uint64))
    np.testing.assert_array_equal(nS, [n_rounds, n_rounds])

    key = pyro.seed_key(seed)
    ctx = pyro.Context.pyrochlore(L, 1, pyro.FP64, seed=key)  # for the colouring only
    ctx.define_J1J2J3(1.0, -0.05, 0.05, 1.0)
    ctx.setup()
    colours, n_col = ctx.colouring(), ctx.info()["n_colours"]
    ctx.close()
    os_ = []
    for rep in range(R):
        o = oracle.Oracle.pyrochlore(L, 1.0, -0.05, 0.05)
        o.randomize_ctr(key, rep)
        os_.append(o)
    p = 0
    for rep in range(R):  # burn-in: 4 sweeps at the starting rung (slot = replica)
        p_end = _oracle_colour_sweeps(os_[rep], rep, T_grid[rep], colours, n_col, key, p, 4)
    p = p_end
    sampled = sorted([int(np.flatnonzero(T_grid == T)[0]) for T in (0.4, 0.25)], key=lambda s: T_grid[s])
    rep_of = np.arange(R, dtype=np.int32)
    att = np.zeros(R, np.uint64)
    accp = np.zeros(R, np.uint64)
    Es, E2s = np.zeros(R), np.zeros(R)
    S = np.zeros((3, 2, L ** 3, 2))
    S2 = np.zeros_like(S)
    M = np.zeros((3, 2, L ** 3, 16, 16, 2))
    swap_seed = pyro.seed_key(seed + R)
    for rnd in range(n_rounds):
        slot_of = np.argsort(rep_of)
        for rep in range(R):
            p_end = _oracle_colour_sweeps(os_[rep], rep, T_grid[slot_of[rep]], colours, n_col, key, p, 2)
        p = p_end
        for j, sl in enumerate(sampled):
            out, m = os_[rep_of[sl]].ssf_sample(("xx", "yy", "zz"), munu=True)
            S[:, j] += out
            S2[:, j] += out ** 2
            M[:, j] += m
        e = np.array([o.energy_per_cell() for o in os_])
        Es += e[rep_of]
        E2s += e[rep_of] ** 2
        rep_of = oracle.exchange(T_grid, e, L ** 3, rep_of, rnd, swap_seed, att, accp)
    scale = 6 + 12 * 0.05 + 12 * 0.05
    np.testing.assert_allclose(E, Es[::-1], rtol=0, atol=1e-12 * 16 * scale * n_rounds)
    np.testing.assert_allclose(E2, E2s[::-1], rtol=0, atol=1e-12 * (16 * scale) ** 2 * n_rounds)
    N = 16 * L ** 3
    np.testing.assert_allclose(Sf, S, rtol=0, atol=1e-12 * N ** 2 * n_rounds)
    np.testing.assert_allclose(S2f, S2, rtol=0, atol=1e-12 * N ** 4 * n_rounds)
    if munu:
        # FP32 MFMA products: ~1e-7 relative to |A|^2 ~ L^3 per term
        assert np.abs(Mf - M).max() <= 2e-5 * L ** 3 * n_rounds
    # per-pair swap summary printed by the driver, decided as the oracle decides
    pairs = re.findall(r"pair\s+(\d+)<->\s*(\d+).*\((\d+)/(\d+)\)", r.stdout)
    assert len(pairs) == R - 1
    for a, b, acc, at in pairs:
        a = int(a)
        assert int(b) == a + 1
        assert int(acc) == accp[a] and int(at) == att[a]
        # checkerboard: pair a is attempted in the rounds of parity a % 2
        assert int(at) == len([k for k in range(n_rounds) if k % 2 == a % 2])
    assert accp.sum() > 0  # the run must see swaps
    with _h5(tmp_path / (stem + ".spins.h5")) as f:
        # rounding of the shell-grouped field sums against the oracle's
        # per-bond sums, accumulated over 36 passes
        np.testing.assert_allclose(f.read("spin_orientation"), os_[rep_of[R - 1]].get_spins(), rtol=0, atol=1e-9)


@pytest.mark.gpu
def test_anneal_seeds_merge(gpu, tmp_path):
    """Two seeds of one anneal merged by acc_runs: one group, summed energy
    and SSF, cross-seed variance of the energy per cell."""
    from pyro_cmc_amd.h5lite import H5File

    base = ["-o", str(tmp_path), "--J1", "1", "--T_sample", "0.3", "--T_hot", "2", "--n_steps", "3",
            "-w", "1", "--n_burn_in", "1", "--n_sample", "2", "--fp32", "2"]
    for seed in ("4", "5"):
        r = _run([_tool(gpu, "anneal"), "-s", seed, *base])
        assert r.returncode == 0, r.stderr
    r = _run([_tool(gpu, "acc_runs"), str(tmp_path)])
    assert r.returncode == 0, r.stderr
    runs = sorted(tmp_path.glob("*.out.h5"))
    merged = sorted(tmp_path.glob("*merge2_*.avg.h5"))
    assert len(runs) == 2 and len(merged) == 1
    E, S = [], []
    for p in runs:
        with _h5(p) as f:
            E.append(f.read("/energy/E"))
            S.append(f.read("/ssf/static_corr"))
    with H5File(merged[0]) as f:
        np.testing.assert_allclose(f.read("/energy/E"), E[0] + E[1], rtol=1e-14)
        np.testing.assert_allclose(f.read("/energy/var"), np.var(np.array(E), axis=0), rtol=1e-9, atol=1e-14)
        np.testing.assert_allclose(f.read("/ssf/static_corr"), S[0] + S[1], rtol=1e-14)
        assert int(f.read("/energy/n_seeds")) == 2


@pytest.mark.gpu
def test_fieldcool_file_matches_oracle_replay(gpu, oracle, tmp_path):
    """fieldcool.cpp:148-205 replayed by the oracle: anneal in a field,
    release it, re-burn at the final temperature, sample the SSF."""
    pyro = gpu
    args = ["-o", str(tmp_path), "-s", "6", "--J1", "1", "--J2", "0.1", "-B", "0", "0", "0.3",
            "--T_hot", "2", "--T_cold", "0.2", "--n_steps", "4", "--n_sweep", "2", "--n_burn_in", "3",
            "--n_sample", "2", "--save_state", "2"]
    r = _run([_tool(pyro, "fieldcool"), *args])
    assert r.returncode == 0, r.stderr
    stem = "run_L=2_J1=1_J2=0.1_J3=0_Delta=1_B=0,0,0.3_seed=6_Tc=0.2_"
    with _h5(tmp_path / (stem + ".out.h5")) as f:
        T_steps = f.read("/energy/T_list")
        E = f.read("/energy/E")
        T_final = f.read("/ssf/T_list")
        S = f.read("/ssf/static_corr")
        S2 = f.read("/ssf/static_corr_2")
        assert f.read("/ssf/corr_lookup") == ["xx", "yy", "zz"]
        np.testing.assert_array_equal(f.read("/ssf/n_samples"), [2])
    f = (0.2 / 2.0) ** (1 / 4)
    np.testing.assert_allclose(T_steps, [2.0 * f ** i for i in range(4)], rtol=1e-14)
    np.testing.assert_allclose(T_final, [2.0 * f ** 4], rtol=1e-14)

    key = pyro.seed_key(6)
    ctx = pyro.Context.pyrochlore(2, 1, pyro.FP64, seed=key)  # for the colouring only
    ctx.define_J1J2J3(1.0, 0.1, 0.0, 1.0)
    ctx.setup()
    colours, n_col = ctx.colouring(), ctx.info()["n_colours"]
    ctx.close()
    o = oracle.Oracle.pyrochlore(2, 1.0, 0.1, 0.0)
    o.set_field(np.array([0, 0, 0.3]))
    o.randomize_ctr(key, 0)
    p = _oracle_colour_sweeps(o, 0, 2.0, colours, n_col, key, 0, 3)
    E_rep = []
    for T in T_steps:
        p = _oracle_colour_sweeps(o, 0, T, colours, n_col, key, p, 2)
        E_rep.append(o.energy_per_cell())
    scale = 6 + 12 * 0.1 + 0.3
    np.testing.assert_allclose(E, E_rep, rtol=0, atol=1e-12 * 16 * scale)
    o.set_field(np.array([0.0, 0.0, 0.0]))
    p = _oracle_colour_sweeps(o, 0, T_final[0], colours, n_col, key, p, 3)
    want, want2 = 0.0, 0.0
    for _ in range(2):
        p = _oracle_colour_sweeps(o, 0, T_final[0], colours, n_col, key, p, 2)
        out = o.ssf_sample(("xx", "yy", "zz"))
        want = want + out
        want2 = want2 + out ** 2
    N = 16 * 8
    np.testing.assert_allclose(S[:, 0], want, rtol=0, atol=1e-12 * N ** 2 * 2)
    np.testing.assert_allclose(S2[:, 0], want2, rtol=0, atol=1e-12 * N ** 4 * 2)
    with _h5(tmp_path / (stem + ".spins.h5")) as f:
        np.testing.assert_allclose(f.read("spin_orientation"), o.get_spins(), rtol=0, atol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [[], ["--fp32"]])
def test_bench_lifted_passes(pyro, extra):
    """Lifted and standard Metropolis on the J1-J2 square lattice agree on
    <E> and <E^2> within 0.1 * min(|J1|, |J2|) (bench_lifted.cpp:254-272),
    at a shorter run than the harness defaults."""
    r = _run([_tool(pyro, "bench_lifted"), "-L", "8", "--n_burn", "500", "--n_sample", "4000", *extra], timeout=300)
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "RESULT: PASS" in r.stdout
    # the table row "<E>  standard  lifted  |difference|" (not the PASS/FAIL line)
    row = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("<E> ") and ":" not in ln][0]
    e_std, e_lft = float(row[1]), float(row[2])
    # ferromagnetic J1 = -1 with J2 = 0.3 (below |J1| / 2: the ferromagnet is
    # the ground state, E/spin = -2 + 2 * 0.3 = -1.4) at T = 0.5: above the
    # ground state, and still well below zero
    assert -1.4 < e_std < -0.4 and -1.4 < e_lft < -0.4


@pytest.mark.gpu
def test_bench_speed_reports_both_variants(pyro):
    r = _run([_tool(pyro, "bench_speed"), "-L", "16", "--n_warm", "20", "--n_sweep", "200", "--reps", "2"], timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert sum(ln.startswith("rep ") for ln in lines) == 2
    ns = [ln for ln in lines if ln.startswith("ns / sweep")][0].split()
    assert float(ns[3]) > 0 and float(ns[4]) > 0
    acc = [ln for ln in lines if ln.startswith("accept")][0].replace("%", " ").split()
    assert 0 < float(acc[1]) <= 100 and 0 < float(acc[2]) <= 100
    assert any(ln.startswith("Faster: ") for ln in lines)
