"""The anneal configurations of the workload list, run at their full sizes
through the batch driver (one context per configuration, one replica per
seed), with checks that do not depend on this implementation:

- plain anneal, FP64, L = 4, J1 = 1 (T 10 -> 0.05 over 200 log steps,
  n_burn_in 16, n_sweep 16, n_sample 128, Metropolis + OR): the classical
  Heisenberg pyrochlore antiferromagnet has E/spin -> -J1 with C/N -> 3/4
  as T -> 0 (Moessner and Chalker 1998), so at T_cold = 0.05 the mean
  E/spin over seeds is -1 + 0.75 * 0.05 up to O(T^2);
- FP64 and FP32 side by side, L = 16, J1 = 1, J2 = 0.1 (T 10 -> 0.02 over
  100 log steps, n_sweep 16): the seed-averaged energy of the two
  precisions agrees within 3 sigma at every step of the anneal;
- the tempering workload, L = 12 x 64 replicas, J = (1, -0.05, 0.05),
  T log-spaced in [0.02, 2], sweeps of 1 Metropolis + 2 OR, exchange every
  4 sweeps, 9-correlator SSF every round, at its full length (4096 burn-in
  and 16384 measured sweeps), six independent ladders per precision: FP32
  against FP64 at every rung for <E>/N, C_v/N and S(q) at three
  wavevectors, within 3 sigma of the seed-to-seed errors, with FP32 in the
  benchmark's fused pass order and FP64 pass by pass;
  and the FP32 run against the reference sampler (the oracle in reference
  mode, 16 independent chains per rung, tests/golden/c3_reference.json) at
  T = 2.0, 0.62 and 0.19;
- the large tempering workload, L = 48 x 16 rungs, J = (1, 0.1, 0.1),
  B = (0, 0, 0.2), T log-spaced in [0.05, 1], sweeps of 1 lifted Metropolis
  + 1 OR, exchange every 2 sweeps, six independent ladders per precision:
  FP32 in the fused pass order (the benchmark's sweep) against FP64 pass by
  pass at every rung for <E>/N, C_v/N and S(q) at three wavevectors, within
  3 sigma of the seed-to-seed errors.  The CPU reference sampler would take
  minutes per sweep of a ladder at this size, so the all-FP64 GPU path is
  the reference here.
"""
import numpy as np
import pytest


def _batch(gpu, L, J, n_seeds, precision, **kw):
    from pyro_cmc_amd import batch

    items = [batch.AnnealItem(L, J) for _ in range(n_seeds)]
    res, _ = batch.run_anneal_batch(items, precision=precision, **kw)
    return res


@pytest.mark.gpu
def test_plain_anneal_fp64_low_temperature_energy(gpu):
    res = _batch(gpu, 4, (1.0, 0.0, 0.0, 1.0), 32, gpu.FP64, T_hot=10.0, T_cold=0.05, n_steps=200,
                 n_sweep=16, n_burn_in=16, n_sample=128, seed=1)
    e = np.array([r.E_final for r in res]) / 16.0  # per spin
    want = -1.0 + 0.75 * 0.05
    sem = e.std(ddof=1) / np.sqrt(len(e))
    # statistical error plus the O(T^2) remainder of the low-T expansion
    assert abs(e.mean() - want) < 3 * sem + 0.05 ** 2, (e.mean(), sem)


@pytest.mark.gpu
def test_fp32_and_fp64_anneals_agree(gpu):
    kw = dict(T_hot=10.0, T_cold=0.02, n_steps=100, n_sweep=16, n_burn_in=16, n_sample=64, seed=2)
    r64 = _batch(gpu, 16, (1.0, 0.1, 0.0, 1.0), 16, gpu.FP64, **kw)
    r32 = _batch(gpu, 16, (1.0, 0.1, 0.0, 1.0), 16, gpu.FP32, **kw)
    e64 = np.array([r.E for r in r64])  # [seed][step], per cell
    e32 = np.array([r.E for r in r32])
    se = np.sqrt(e64.var(0, ddof=1) / len(e64) + e32.var(0, ddof=1) / len(e32))
    z = np.abs(e64.mean(0) - e32.mean(0)) / se
    # 100 steps: allow the expected few 2-sigma steps, none beyond 4 sigma
    assert np.mean(z > 3) < 0.05 and z.max() < 4.5, (z.max(), np.argmax(z))
    f64 = np.array([r.E_final for r in r64]) / 16
    f32 = np.array([r.E_final for r in r32]) / 16
    se_f = np.sqrt(f64.var(ddof=1) / len(f64) + f32.var(ddof=1) / len(f32))
    assert abs(f64.mean() - f32.mean()) < 3 * se_f


N_COPIES = 6


def _tempering_copies(gpu, precision, seed, n_copies=N_COPIES, n_burn=1024, n_rounds=4096, fused=False):
    """The tempering workload at its full length (4096 burn-in and 16384
    measured sweeps) for n_copies independent ladders at once: one context
    holds the copies' 64 rungs back to back, so the only pairs joining two
    copies sit between T = 0.02 and T = 2.0, where a swap needs
    exp(-(1/0.02 - 1/2) (E_hot - E_cold) L^3) ~ exp(-3e4) and never happens
    (checked).  fused: the sweeps in the fused pass order (per colour each
    site's Metropolis and first OR update from one field evaluation, then
    the second OR pass), as the benchmark runs them.  Returns T[64] and per
    copy and rung <E>/cell, <E^2>/cell^2 and S_xx + S_yy + S_zz at
    K = (0, 0, 0), (0, 0, 1), (0, 1, 1)."""
    L, R0 = 12, 64
    R = R0 * n_copies
    T = np.geomspace(2.0, 0.02, R0)
    ctx = gpu.Context.pyrochlore(L, n_replicas=R, precision=precision, seed=seed)
    ctx.define_J1J2J3(1.0, -0.05, 0.05)
    ctx.setup()
    ctx.set_fused_passes(fused)
    ctx.randomize(seed + 1)
    ctx.pt_init(np.tile(T, n_copies), seed + 2)
    ctx.ssf_configure(("xx", "yy", "zz", "xy", "xz", "yz", "yx", "zx", "zy"), n_slots=R, store_error_term=False)
    ctx.pt_rounds(gpu.METROPOLIS, 2, 4, True, n_burn)
    ctx.pt_reset_stats()
    ctx.ssf_reset()
    ctx.pt_rounds(gpu.METROPOLIS, 2, 4, True, n_rounds)
    st = ctx.pt_stats()
    assert np.all(st["n"] == n_rounds)
    boundary = np.arange(R0 - 1, R - 1, R0)
    assert np.all(st["accepts"][boundary] == 0) and np.all(st["attempts"][boundary] > 0)
    ks = [0, 1, L + 1]  # flat K = (K0 L + K1) L + K2: (0, 0, 0), (0, 0, 1), (0, 1, 1)
    S = []
    for slot in range(R):
        Ss, _, n = ctx.ssf_read(slot, with_sq=False)
        assert n == n_rounds
        S.append(Ss[:3, ks, 0].sum(axis=0) / n)
    ctx.close()
    shape = (n_copies, R0)
    return (T, (st["E"] / n_rounds).reshape(shape), (st["E2"] / n_rounds).reshape(shape),
            np.array(S).reshape(n_copies, R0, 3), L ** 3)


def _stats(T, E, E2, S, nc):
    """(mean, seed-to-seed standard error) over copies of <E>/N, C_v/N and
    S(q)/N at the three K, per rung (acc_runs.py:315-331)."""
    N = 16 * nc
    e = E / 16.0  # per spin
    cv = nc * (E2 - E ** 2) / T[None, :] ** 2 / 16.0  # C_v/N: cells * var(E/cell) / T^2 / 16
    s = S / N  # S(q)/N
    out = []
    for x in (e, cv, *(s[:, :, j] for j in range(s.shape[2]))):
        out.append((x.mean(axis=0), x.std(axis=0, ddof=1) / np.sqrt(x.shape[0])))
    return out


@pytest.fixture(scope="module")
def headline_runs(gpu):
    # FP32 in the benchmark's fused pass order, FP64 pass by pass (the
    # reference's order): the comparison covers the precision and the order
    a = _tempering_copies(gpu, gpu.FP32, 21, fused=True)
    b = _tempering_copies(gpu, gpu.FP64, 41, fused=False)
    return a, b


@pytest.mark.gpu
def test_headline_tempering_fp32_matches_fp64(headline_runs):
    """FP32 (fused pass order) against FP64 (pass by pass) at every rung,
    each side's error the spread of
    its independent copies: at the cold end no swap reaches a rung, and a
    single long run's batch means miss the run-to-run spread of its slow
    relaxation (at T = 0.02 two runs differ by several batch-mean errors)."""
    (T, E32, E232, S32, nc), (_, E64, E264, S64, _) = headline_runs
    a, b = _stats(T, E32, E232, S32, nc), _stats(T, E64, E264, S64, nc)
    z = np.array([np.abs(m1 - m2) / np.sqrt(s1 ** 2 + s2 ** 2 + 1e-300) for (m1, s1), (m2, s2) in zip(a, b)])
    assert np.all(np.isfinite(z))
    worst = np.unravel_index(np.argmax(z), z.shape)
    print(f"FP32 vs FP64, {z.size} comparisons over all {len(T)} rungs: {int((z > 3).sum())} beyond 3 sigma, "
          f"largest {z.max():.2f} (observable {worst[0]}, T = {T[worst[1]]:.3g}); per observable beyond 3: "
          f"{(z > 3).sum(1).tolist()}")
    for name, (m1, s1), (m2, s2) in zip(("E/N", "Cv/N", "S0/N", "S1/N", "S2/N"), a, b):
        for i in (0, 16, 32, 48, 63):
            print(f"  T = {T[i]:.4g} {name}: FP32 {m1[i]:.6g}({s1[i]:.2g}) FP64 {m2[i]:.6g}({s2[i]:.2g})")
    # 6 copies per side: z is close to Student's t with about 10 degrees of
    # freedom: P(|t| > 3) = 1.3%, about 4 of the 320 comparisons; 5% would
    # be 16; P(|t| > 6) = 1.3e-4
    assert np.mean(z > 3) < 0.05, (np.sort(z.ravel())[-8:], z.size)
    assert z.max() < 6, np.sort(z.ravel())[-8:]


@pytest.mark.gpu
def test_headline_tempering_matches_reference_sampler(headline_runs):
    """FP32 tempering at three rungs against independent chains of the
    reference sampler at the same temperature, both with seed-to-seed
    errors (acc_runs.py:315-331): <E>/N, C_v/N and S(q)/N within 3 sigma."""
    import json
    from pathlib import Path

    ref = json.loads((Path(__file__).parent / "golden" / "c3_reference.json").read_text())
    T, E, E2, S, nc = headline_runs[0]
    st = _stats(T, E, E2, S, nc)
    zs = []
    for rung in ref["rungs"]:
        i = rung["rung"]
        assert abs(T[i] - rung["T"]) < 1e-12
        refs = [rung["e_per_spin"], rung["cv_per_spin"], *rung["S_per_N"]]
        for name, (m, s), (rm, rs) in zip(("E/N", "Cv/N", "S0/N", "S1/N", "S2/N"), st, refs):
            z = abs(m[i] - rm) / np.hypot(s[i], rs)
            zs.append(z)
            print(f"T = {T[i]:.4g} {name}: GPU {m[i]:.6g}({s[i]:.2g}) reference {rm:.6g}({rs:.2g}) z = {z:.2f}")
    zs = np.array(zs)
    # 15 comparisons: one beyond 3 sigma is a 2-3% event; none beyond 4
    assert np.sum(zs > 3) <= 1 and zs.max() < 4, zs


def _large_tempering_copies(gpu, precision, seed, fused, n_copies=N_COPIES, n_burn=64, n_rounds=192):
    """The large tempering workload for n_copies independent ladders stacked
    in one context: L = 48, 16 rungs per ladder at T log-spaced in
    [0.05, 1], J = (1, 0.1, 0.1), B = (0, 0, 0.2), rounds of 2 sweeps (1
    lifted Metropolis + 1 OR each) with an exchange and one SSF sample per
    round.  The pairs joining two ladders sit between T = 0.05 and T = 1,
    where a swap needs exp(-19 (E_hot - E_cold) L^3) and never happens
    (checked).  Returns T[16] and per ladder and rung <E>/cell, <E^2>/cell^2
    and S_xx + S_yy + S_zz at K = (0, 0, 0), (0, 0, 1), (0, 1, 1)."""
    L, R0 = 48, 16
    R = R0 * n_copies
    T = np.geomspace(1.0, 0.05, R0)
    ctx = gpu.Context.pyrochlore(L, n_replicas=R, precision=precision, seed=seed)
    ctx.define_J1J2J3(1.0, 0.1, 0.1)
    ctx.setup()
    ctx.set_fused_passes(fused)
    ctx.set_field(np.array([0.0, 0.0, 0.2]))
    ctx.randomize(seed + 1)
    ctx.pt_init(np.tile(T, n_copies), seed + 2)
    ctx.ssf_configure(("xx", "yy", "zz"), n_slots=R, store_error_term=False)
    ctx.pt_rounds(gpu.LIFTED, 1, 2, True, n_burn)
    ctx.pt_reset_stats()
    ctx.ssf_reset()
    ctx.pt_rounds(gpu.LIFTED, 1, 2, True, n_rounds)
    st = ctx.pt_stats()
    assert np.all(st["n"] == n_rounds)
    boundary = np.arange(R0 - 1, R - 1, R0)
    assert np.all(st["accepts"][boundary] == 0) and np.all(st["attempts"][boundary] > 0)
    ks = [0, 1, L + 1]
    S = []
    for slot in range(R):
        Ss, _, n = ctx.ssf_read(slot, with_sq=False)
        assert n == n_rounds
        S.append(Ss[:3, ks, 0].sum(axis=0) / n)
    ctx.close()
    shape = (n_copies, R0)
    return (T, (st["E"] / n_rounds).reshape(shape), (st["E2"] / n_rounds).reshape(shape),
            np.array(S).reshape(n_copies, R0, 3), L ** 3)


@pytest.mark.gpu
def test_large_tempering_fp32_matches_fp64(gpu):
    """FP32 (fused pass order, as the benchmark runs it) against FP64 (pass
    by pass) at all 16 rungs of the large tempering workload, each side's
    error the spread of its independent ladders.  384 measured sweeps do
    not equilibrate the coldest rungs of a 1.8-million-spin lattice: there
    E / N still drifts during the measurement and the "C_v" of the run is
    mostly that drift, which the two precisions must then show alike."""
    (T, E32, E232, S32, nc) = _large_tempering_copies(gpu, gpu.FP32, 61, True)
    (_, E64, E264, S64, _) = _large_tempering_copies(gpu, gpu.FP64, 81, False)
    a, b = _stats(T, E32, E232, S32, nc), _stats(T, E64, E264, S64, nc)
    z = np.array([np.abs(m1 - m2) / np.sqrt(s1 ** 2 + s2 ** 2 + 1e-300) for (m1, s1), (m2, s2) in zip(a, b)])
    assert np.all(np.isfinite(z))
    worst = np.unravel_index(np.argmax(z), z.shape)
    print(f"L=48 FP32 vs FP64, {z.size} comparisons over all {len(T)} rungs: {int((z > 3).sum())} beyond 3 "
          f"sigma, largest {z.max():.2f} (observable {worst[0]}, T = {T[worst[1]]:.3g})")
    for name, (m1, s1), (m2, s2) in zip(("E/N", "Cv/N", "S0/N", "S1/N", "S2/N"), a, b):
        for i in (0, 8, 15):
            print(f"  T = {T[i]:.4g} {name}: FP32 {m1[i]:.6g}({s1[i]:.2g}) FP64 {m2[i]:.6g}({s2[i]:.2g})")
    # 6 ladders per side (t with about 10 degrees of freedom): 80
    # comparisons, about one beyond 3 sigma expected
    assert np.mean(z > 3) < 0.05, (np.sort(z.ravel())[-8:], z.size)
    assert z.max() < 6, np.sort(z.ravel())[-8:]
