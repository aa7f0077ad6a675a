"""The large tempering run and the uneven batch at their own workloads.

Large tempering run: L = 48 (1,769,472 spins) x 16 replicas at log-spaced T
in [0.05, 1], J = (1, 0.1, 0.1), B = (0, 0, 0.2), each sweep one lifted
Metropolis + one OR pass on the default 6x6x4 tiles, an exchange every 2
sweeps, the 9 correlators with the sublattice-resolved matrices every sweep
through the 192-sample FP32-MFMA buffer.
  - one sweep of two replicas replayed by the oracle in the exported tiled
    order, pass by pass and with fused passes: FP64 decisions exact, FP32
    within the first-order error bound;
  - tempering rounds: the sampled E sums equal the energies of the states
    the rounds leave, and the accumulated S(q) and C_{mu nu}(K) equal sums
    over rounds of products of the spins' own transforms, at a sample of K
    and for the first and last slot (the largest buffer offsets).

Uneven batch: 256 anneals, L drawn from {4, 8, 12, 16, 24}, J2 and J3 in
[-0.2, 0.2], |B| in [0, 0.5] along a random direction, all in one job: every
item's final energy equals the oracle's energy of its returned spins under
its own model.
"""
import numpy as np
import pytest

import fp32_bound

pytestmark = pytest.mark.gpu

L5, R5 = 48, 16
J5 = (1.0, 0.1, 0.1)
B5 = (0.0, 0.0, 0.2)
T5 = np.geomspace(1.0, 0.05, R5)
CORR9 = ("xx", "yy", "zz", "xy", "xz", "yz", "yx", "zx", "zy")


def _scale(J, B):
    return 6 * abs(J[0]) + 12 * abs(J[1]) + 12 * abs(J[2]) + float(np.linalg.norm(B))


@pytest.fixture(scope="module")
def oracle48(oracle):
    o = oracle.Oracle.pyrochlore(L5, *J5)
    o.set_field(np.array(B5))
    return o


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("fused", [False, True])
def test_large_tempering_sweep_replay(gpu, oracle48, prec, fused):
    seed = 0x48C5
    ctx = gpu.Context.pyrochlore(L5, n_replicas=R5, precision=prec, seed=seed)
    ctx.define_J1J2J3(*J5)
    ctx.setup()
    ctx.set_fused_passes(fused)
    ctx.set_field(np.array(B5))
    ctx.set_temperatures(T5)
    ctx.randomize(17)
    if prec == 64:
        try:
            ctx.set_tile_shape((6, 6, 4))
        except gpu.PyroError as e:
            assert e.code == 5  # FP64 tile + halo beyond this device's LDS
            ctx.set_tile_shape(None)
    assert prec == 64 or ctx.tile_shape() == (6, 6, 4)
    site, koff = ctx.sweep_order(1)
    reps = (0, R5 - 1) if prec == 64 else (0,)
    before = {r: (ctx.get_spins(r), ctx.get_lifted_dir(r)) for r in reps}
    p0 = ctx.pass_count()
    acc = ctx.sweep(gpu.LIFTED, 1, 1)
    o = oracle48
    for r in reps:
        o.set_spins(before[r][0])
        o.set_dir(before[r][1])
        if prec == 64:
            a = o.sweep_order(gpu.LIFTED, T5[r], site, koff, seed, r, p0, shared=fused)
            assert int(acc[r]) == a
            assert np.abs(ctx.get_spins(r) - o.get_spins()).max() < 1e-9
            assert np.array_equal(ctx.get_lifted_dir(r), o.get_dir())
        else:
            info = np.empty((len(site), 4))
            o.sweep_order(gpu.LIFTED, T5[r], site, koff, seed, r, p0, info=info, shared=fused)
            adj = fp32_bound.padded_adjacency(o, (*J5, 1.0))
            E, taint, seeds = fp32_bound.propagate(site, fp32_bound.tiled_stages(ctx, site, koff, fused), info, adj,
                                                   T5[r], _scale(J5, B5), ctx.N)
            fp32_bound.check(ctx.get_spins(r), o.get_spins(), E, taint, seeds, len(site),
                             f"L=48 replica {r} T={T5[r]:.3g}")
    ctx.close()


def test_large_tempering_rounds_energy_and_munu(gpu):
    """Tempering rounds of the large run with mu-nu sampling, checked
    against quantities read from the spins after every round."""
    ctx = gpu.Context.pyrochlore(L5, n_replicas=R5, precision=gpu.FP32, seed=5)
    ctx.define_J1J2J3(*J5)
    ctx.setup()
    ctx.set_field(np.array(B5))
    ctx.randomize(6)
    ctx.pt_init(T5, 7)
    ctx.pt_exchange_interval(2)
    ctx.ssf_configure(CORR9, n_slots=R5, munu=True, munu_buffer=192)
    ctx.pt_rounds(gpu.LIFTED, 1, 1, False, 4)  # a few sweeps off the random start
    ctx.pt_reset_stats()
    n_k = L5 ** 3
    rng = np.random.default_rng(3)
    ks = np.concatenate([[0, 1, L5, L5 * L5 + L5 + 1], rng.choice(n_k, 60, replace=False)])
    pos = gpu.pyrochlore_cell()[0].astype(float)
    Kv = np.stack([ks // (L5 * L5), (ks // L5) % L5, ks % L5], 1)
    q = 2 * np.pi * np.where(2 * Kv < L5, Kv, Kv - L5) / (8.0 * L5)
    ph = np.exp(-1j * q @ pos.T)  # [k][mu] e^{-i q.r_mu}
    ax = {"x": 0, "y": 1, "z": 2}
    al = [ax[c[0]] for c in CORR9]
    be = [ax[c[1]] for c in CORR9]
    slots = (0, R5 - 1)
    want_E = np.zeros(R5)
    want_M = {s: np.zeros((len(CORR9), len(ks), 16, 16), complex) for s in slots}
    want_S = {s: np.zeros((len(CORR9), len(ks)), complex) for s in slots}
    n_rounds = 6
    for _ in range(n_rounds):
        rep_of = ctx.pt_permutation()  # slots at sampling time (the round's exchange comes last)
        ctx.pt_rounds(gpu.LIFTED, 1, 1, True, 1)
        e = ctx.energy_per_cell()
        want_E += e[rep_of]
        for s in slots:
            ft = ctx.spin_ft(int(rep_of[s]))  # [3][16][n_k][2]
            A = ft[:, :, ks, 0] + 1j * ft[:, :, ks, 1]  # [a][mu][k]
            b = np.einsum("km,amk->ak", ph, A)  # b^a(K) = sum_mu e^{-iq.r_mu} A^a_mu(K)
            for c in range(len(CORR9)):
                want_M[s][c] += np.conj(A[al[c]]).T[:, :, None] * A[be[c]].T[:, None, :]
                want_S[s][c] += np.conj(b[al[c]]) * b[be[c]]
    st = ctx.pt_stats()
    assert np.all(st["n"] == n_rounds)
    tol_e = 1e-5 * 16 * _scale(J5, B5)
    assert np.allclose(st["E"], want_E, rtol=0, atol=tol_e * n_rounds), np.abs(st["E"] - want_E).max()
    N = ctx.N
    for s in slots:
        S, _, n = ctx.ssf_read(s)
        assert n == n_rounds
        got = S[:, ks, 0] + 1j * S[:, ks, 1]
        assert np.abs(got - want_S[s]).max() <= 1e-5 * N ** 2 * n_rounds / 16
        for c in (0, 7):  # xx (an upper tile) and zx (a lower tile, read by conjugation)
            m = ctx.ssf_read_munu(s, c)[ks]
            got = m[..., 0] + 1j * m[..., 1]
            # FP32 products: error ~ 1e-7 relative to |A|^2 ~ L^3 per term
            assert np.abs(got - want_M[s][c]).max() <= 2e-5 * n_k * n_rounds
    ctx.close()


@pytest.mark.parametrize("fused", [False, True])
def test_uneven_batch_every_item_against_its_oracle(gpu, oracle, fused):
    """All 256 items of the uneven batch, pass by pass and in the fused pass
    order the benchmark runs: each item's final energy against an oracle of
    its own model."""
    from pyro_cmc_amd import batch

    items = batch.uneven_items(256)
    assert sorted({it.L for it in items}) == [4, 8, 12, 16, 24]
    res, stats = batch.run_anneal_batch(items, n_steps=8, n_sweep=4, n_burn_in=4, n_sample=2, keep_spins=True,
                                        seed=11, fused=fused)
    assert stats["contexts"] == {L: sum(it.L == L for it in items) for L in (4, 8, 12, 16, 24)}
    orcs = {}
    worst = 0.0
    for it, r in zip(items, res):
        o = orcs.get(it.L)
        if o is None:
            o = orcs[it.L] = oracle.Oracle.pyrochlore(it.L, 1.0, 1.0, 1.0)
        o.set_J1J2J3(*it.J)
        o.set_field(np.array(it.B))
        o.set_spins(r.spins)
        assert np.allclose(np.linalg.norm(r.spins, axis=1), 1.0, atol=1e-5)
        d = abs(r.E_final - o.energy_per_cell()) / (16 * _scale(it.J, it.B))
        worst = max(worst, d)
        assert d <= 1e-5, (it, r.E_final, o.energy_per_cell())
        assert r.E.shape == (len(r.T_list),) and np.all(np.isfinite(r.E))
        assert np.array_equal(r.n_samples, [2])
    print(f"uneven batch: worst relative energy deviation {worst:.3g} over {len(items)} items")
