"""Batches of independent models in one context, and device-side anneal
schedules.

Every replica may carry its own pyrochlore couplings (J1, J2, J3, Delta) and
field; each replica is checked against an oracle instance built with that
replica's model (fields and energy to 1e-12 / 1e-5, FP64 trajectories draw
for draw in colour order and in the tiled order).  The anneal schedule
(temperature steps run as a replayed HIP graph) must equal the same steps
issued one call at a time, and contexts of different sizes run concurrently
on their own streams must give what each gives alone, and a Batch of
contexts (one graph per step, the small lattices in one persistent launch)
what each gives with its own steps.
"""
import numpy as np
import pytest

from helpers import random_spins

pytestmark = pytest.mark.gpu

MODELS = [  # (J1, J2, J3, Delta), B
    ((1.0, -0.2, 0.15, 1.0), (0.0, 0.0, 0.0)),
    ((-1.0, 0.1, -0.05, 0.7), (0.1, -0.3, 0.2)),
    ((0.5, 0.0, 0.0, 1.3), (0.0, 0.0, 0.4)),
    ((1.0, 0.2, 0.2, 1.0), (-0.2, 0.1, 0.0)),
]
TOL = {64: 1e-12, 32: 1e-5}


def _batch(pyro, oracle, L, prec, seed=4321, context_J=(1.0, 0.1, 0.1)):
    R = len(MODELS)
    ctx = pyro.Context.pyrochlore(L, n_replicas=R, precision=prec, seed=seed)
    ctx.define_J1J2J3(*context_J)
    ctx.setup()
    ctx.set_replica_couplings(np.array([m[0] for m in MODELS]))
    os_ = []
    for r, (J, B) in enumerate(MODELS):
        ctx.set_field(np.array(B), r)
        o = oracle.Oracle.pyrochlore(L, *J)
        o.set_field(np.array(B))
        os_.append(o)
    return ctx, os_


def _scale(J, B):
    return 6 * abs(J[0]) * max(1.0, abs(J[3])) + 12 * abs(J[1]) + 12 * abs(J[2]) + np.linalg.norm(B)


@pytest.mark.parametrize("prec", [64, 32])
def test_replica_couplings_fields_and_energy(gpu, oracle, prec):
    L = 3
    ctx, os_ = _batch(gpu, oracle, L, prec)
    rng = np.random.default_rng(7)
    for r, o in enumerate(os_):
        S = random_spins(ctx.N, rng)
        ctx.set_spins(S, r)
        o.set_spins(S)
    e = ctx.energy_per_cell()
    for r, ((J, B), o) in enumerate(zip(MODELS, os_)):
        s = _scale(J, B)
        assert np.max(np.abs(ctx.local_fields(r) - o.local_fields())) <= TOL[prec] * s
        assert abs(e[r] - o.energy_per_cell()) <= TOL[prec] * 16 * s


def _oracle_sweeps(o, r, kind, n_or, n_sweeps, T, colours, n_col, seed, pass0):
    acc, p = 0, pass0
    for _ in range(n_sweeps):
        acc += o.pass_ctr(kind, T, colours, n_col, seed, r, p)
        p += 1
        for _ in range(n_or):
            o.pass_ctr(3, T, colours, n_col, seed, r, p)
            p += 1
    return acc


@pytest.mark.parametrize("strategy", [3, 4])
@pytest.mark.parametrize("kind,n_or", [(0, 1), (1, 1), (2, 0)])
def test_replica_couplings_trajectory(gpu, oracle, strategy, kind, n_or):
    L, seed = 2, 0xABCDEF
    T = np.array([0.3, 0.9, 2.0, 0.5])
    ctx, os_ = _batch(gpu, oracle, L, 64, seed=seed)
    ctx.set_strategy(strategy)
    ctx.set_temperatures(T)
    rng = np.random.default_rng(11)
    for r, o in enumerate(os_):
        S = random_spins(ctx.N, rng)
        ctx.set_spins(S, r)
        o.set_spins(S)
    colours = ctx.colouring()
    n_col = ctx.info()["n_colours"]
    acc = ctx.sweep(kind, n_or, 3)
    for r, o in enumerate(os_):
        assert int(acc[r]) == _oracle_sweeps(o, r, kind, n_or, 3, T[r], colours, n_col, seed, 0)
        assert np.max(np.abs(ctx.get_spins(r) - o.get_spins())) < 1e-12
        assert np.array_equal(ctx.get_lifted_dir(r), o.get_dir())


@pytest.mark.parametrize("kind,n_or", [(0, 2), (1, 1)])
def test_replica_couplings_tiled(gpu, oracle, kind, n_or):
    L, seed = 8, 0x5EED
    T = [0.4, 1.1, 0.7, 2.5]
    ctx, os_ = _batch(gpu, oracle, L, 64, seed=seed)
    ctx.set_strategy(5)
    ctx.set_temperatures(T)
    rng = np.random.default_rng(13)
    for r, o in enumerate(os_):
        S = random_spins(ctx.N, rng)
        ctx.set_spins(S, r)
        o.set_spins(S)
    site, koff = ctx.sweep_order(n_or)
    try:
        acc = ctx.sweep(kind, n_or, 1)
    except gpu.PyroError as e:
        assert e.code == 5
        pytest.skip(str(e))
    for r, o in enumerate(os_):
        assert int(acc[r]) == o.sweep_order(kind, T[r], site, koff, seed, r, 0)
        assert np.abs(ctx.get_spins(r) - o.get_spins()).max() < 1e-10


def test_replica_couplings_validation(gpu, oracle):
    ctx = gpu.Context.pyrochlore(2, n_replicas=2, precision=32, seed=1)
    ctx.define_J1J2J3(1.0, 0.1, 0.0)  # no J3 shell: 4-colour schedule
    ctx.setup()
    with pytest.raises(gpu.PyroError) as e:
        ctx.set_replica_couplings([[1, 0.1, 0.05, 1], [1, 0, 0, 1]])
    assert e.value.code == 1
    ctx.set_replica_couplings([[1, 0.1, 0.0, 1], [-1, 0.0, 0.0, 0.5]])
    ctx.set_strategy(1)
    ctx.set_temperatures([1.0, 1.0])
    with pytest.raises(gpu.PyroError) as e:
        ctx.sweep()
    assert e.value.code == 5
    g = gpu.Context.pyrochlore(2, n_replicas=2, precision=32, seed=1)
    for name, rel, J in oracle.j1j2j3_couplings(1.0, 0.0, 0.0):
        g.define_coupling(name, rel, J)
    g.setup()
    with pytest.raises(gpu.PyroError) as e:
        g.set_replica_couplings([[1, 0, 0, 1], [1, 0, 0, 1]])
    assert e.value.code == 5


def _anneal_eager(pyro, ctx, sched, kind, n_or, n_sweep, ssf_from):
    E = []
    for k, T in enumerate(sched):
        ctx.set_temperatures(np.full(ctx.R, T))
        ctx.sweep(kind, n_or, n_sweep, count=False)
        if k >= ssf_from:
            ctx.ssf_sample()
        E.append(ctx.energy_per_cell())
    return np.array(E).T


@pytest.mark.parametrize("prec,strategy,fused", [(32, 3, False), (64, 4, False), (32, 5, False), (32, 3, True),
                                                (64, 4, True), (32, 5, True)])
def test_anneal_steps_equal_eager_steps(gpu, oracle, prec, strategy, fused):
    L = 8 if strategy == 5 else 3
    sched = np.geomspace(3.0, 0.1, 9)
    out = []
    for graph in (True, False):
        ctx, _ = _batch(gpu, oracle, L, prec, seed=77)
        ctx.set_strategy(strategy)
        ctx.set_fused_passes(fused)
        ctx.randomize(5)
        ctx.ssf_configure(("xx", "zz", "xy"), n_slots=ctx.R)
        if graph:
            ctx.anneal_schedule(sched)
            ctx.anneal_steps(gpu.LIFTED, 1, 2, False, 6)
            ctx.anneal_steps(gpu.LIFTED, 1, 2, True, 3)
            E = ctx.anneal_history()
        else:
            E = _anneal_eager(gpu, ctx, sched, gpu.LIFTED, 1, 2, 6)
        out.append((E, [ctx.get_spins(r) for r in range(ctx.R)], [ctx.ssf_read(r)[0] for r in range(ctx.R)],
                    ctx.pass_count()))
    (Eg, Sg, Qg, pg), (Ee, Se, Qe, pe) = out
    assert pg == pe
    if strategy in (3, 4):
        # stencil strategies: the step's last pass sums the energy (bonds to
        # earlier colours), the eager steps use the energy kernel
        assert np.allclose(Eg, Ee, rtol=0, atol=1e-5 * 16 * 8)
    else:
        assert np.array_equal(Eg, Ee)
    for a, b in zip(Sg, Se):
        assert np.array_equal(a, b)
    for a, b in zip(Qg, Qe):
        assert np.array_equal(a, b)


def test_concurrent_contexts_match_sequential(gpu, oracle):
    """Contexts of different L enqueued in turn run concurrently on their
    own streams; each ends where it ends when run alone."""
    sched = np.geomspace(5.0, 0.2, 12)
    sizes = (2, 4, 8)

    def run(interleave):
        ctxs = []
        for i, L in enumerate(sizes):
            c, _ = _batch(gpu, oracle, L, 32, seed=100 + i)
            c.randomize(200 + i)
            c.anneal_schedule(sched)
            ctxs.append(c)
        if interleave:
            for _ in range(4):
                for c in ctxs:
                    c.anneal_steps(gpu.METROPOLIS, 2, 3, False, 3)
        else:
            for c in ctxs:
                c.anneal_steps(gpu.METROPOLIS, 2, 3, False, 12)
                c.synchronize()
        return [(c.anneal_history(), c.get_spins(1)) for c in ctxs]

    a, b = run(True), run(False)
    for (Ea, Sa), (Eb, Sb) in zip(a, b):
        assert np.array_equal(Ea, Eb)
        assert np.array_equal(Sa, Sb)


def test_uneven_batch_driver(gpu, oracle):
    """Items of different L, couplings and fields in one job: each result
    belongs to its item (final energy of the returned spins under the item's
    own model equals the last recorded energy), on the anneal grid."""
    from pyro_cmc_amd import batch

    items = batch.uneven_items(n=9, sizes=(2, 4, 8), seed=3)
    res, stats = batch.run_anneal_batch(items, T_hot=5.0, T_cold=0.1, n_steps=6, n_sweep=2, n_burn_in=2,
                                        n_sample=3, keep_spins=True, seed=1)
    grid, idx = batch.generate_T_profile(5.0, 0.1, [0.1], 6)
    assert len(grid) == 7 and idx == [6] and np.all(np.diff(grid) < 0)
    assert stats["contexts"] == {L: sum(it.L == L for it in items) for L in (2, 4, 8) if any(it.L == L for it in items)}
    for it, r in zip(items, res):
        assert r.item is it
        assert np.array_equal(r.T_list, grid) and r.E.shape == (7,) and np.array_equal(r.n_samples, [3])
        o = oracle.Oracle.pyrochlore(it.L, *it.J)
        o.set_field(np.array(it.B))
        o.set_spins(r.spins)
        e = o.energy_per_cell()
        assert abs(r.E_final - e) <= 1e-5 * 16 * _scale(it.J, it.B)
        assert np.all(np.isfinite(r.S)) and r.S.shape == (1, 3, it.L ** 3, 2)


@pytest.mark.parametrize("prec,fused,kind", [(32, True, 0), (32, False, 1), (64, True, 1), (64, False, 0)])
def test_batch_launch_equals_own_graphs(gpu, oracle, prec, fused, kind):
    """A Batch steps every context as its own anneal steps would, bit for
    bit: the L = 2, 4, 6 contexts sweep in the one persistent launch (the
    per-colour stencil order), the L = 16 context (tiled, a large lattice)
    in its own branch of the step's graph.  Energy histories, final spins and
    SSF sums at two sampling temperatures are identical."""
    from pyro_cmc_amd import batch

    items = batch.uneven_items(n=12, sizes=(2, 4, 6, 16), seed=2)
    kw = dict(T_hot=5.0, T_cold=0.1, n_steps=5, n_sweep=2, n_burn_in=2, n_sample=2, seed=2, keep_spins=True,
              precision=prec, fused=fused, kind=kind, T_sample=[1.0, 0.1])
    a, sa = batch.run_anneal_batch(items, persistent=True, **kw)
    b, sb = batch.run_anneal_batch(items, persistent=False, **kw)
    assert sa["contexts"] == {2: 3, 4: 5, 6: 2, 16: 2}
    launch = sa["persistent_launch"]
    assert launch["contexts"] == 3 and launch["blocks_per_workgroup"] == 1
    # replicas placed whole on XCDs: 3 + 5 (one block each) and 2 x 4 blocks
    assert launch["workgroups"] % 8 == 0 and launch["workgroups"] >= 3 + 5 + 2 * 4
    for ra, rb in zip(a, b):
        assert np.array_equal(ra.E, rb.E)
        assert np.array_equal(ra.spins, rb.spins)
        assert np.array_equal(ra.S, rb.S) and np.array_equal(ra.S2, rb.S2)
        assert np.array_equal(ra.n_samples, rb.n_samples) and np.array_equal(ra.n_samples, [2, 2])


def test_batch_takes_forced_xcd_members_into_its_launch(gpu, oracle):
    """A member forced to the XCD-resident sweep (strategy 6) sweeps in the
    batch's persistent launch, whose results it reproduces, rather than as
    a second persistent kernel beside it; a member forced to per-colour
    launches (3) keeps its own.  Both end where their own steps end."""
    sched = np.geomspace(3.0, 0.2, 6)

    def build():
        cs = []
        for i, (L, strat) in enumerate(((4, 6), (6, 3), (3, 0))):
            c, _ = _batch(gpu, oracle, L, 32, seed=70 + i)
            c.set_strategy(strat)
            c.randomize(30 + i)
            c.anneal_schedule(sched)
            cs.append(c)
        return cs

    a = build()
    b = gpu.Batch(a)
    b.anneal_steps(gpu.METROPOLIS, 1, 2, False, 6)
    assert b.info()[0] == 2  # the strategy-6 and the automatic member
    own = build()
    for c in own:
        c.anneal_steps(gpu.METROPOLIS, 1, 2, False, 6)
    for x, y in zip(a, own):
        assert np.array_equal(x.anneal_history(), y.anneal_history())
        for r in range(x.R):
            assert np.array_equal(x.get_spins(r), y.get_spins(r))
    b.close()
    for c in a + own:
        c.close()


def test_batch_wait_timeout_is_reported(gpu, oracle):
    """The persistent launch bounds each member's stage waits by that
    member's own bound: with one member's set to a nanosecond its next read
    reports the timeout, the other member reads normally, and once the bound
    is restored the batch steps and reads again."""
    sched = np.geomspace(3.0, 0.2, 8)
    ctxs = []
    for i, L in enumerate((4, 6)):
        c, _ = _batch(gpu, oracle, L, 32, seed=50 + i)
        c.randomize(9 + i)
        c.anneal_schedule(sched)
        ctxs.append(c)
    b = gpu.Batch(ctxs)
    ctxs[0].set_wait_timeout(1e-9)
    b.anneal_steps(gpu.METROPOLIS, 1, 3, False, 2)
    assert b.info()[0] == 2
    with pytest.raises(gpu.PyroError, match="timed out"):
        ctxs[0].anneal_history()
    assert np.all(np.isfinite(ctxs[1].anneal_history()[:, :2]))
    ctxs[0].set_wait_timeout(0.25)
    b.anneal_steps(gpu.METROPOLIS, 1, 3, False, 2)
    E = ctxs[0].anneal_history()  # (steps 0 and 1 summed partials the aborted launch left unwritten)
    assert np.all(np.isfinite(E[:, 2:4])) and np.all(E[:, 2:4] != 0)
    b.close()
    for c in ctxs:
        c.close()


def test_batch_samples_every_T_sample(gpu, oracle):
    """With several sampling temperatures (C2's {1.0, 0.2, 0.05}) the batch
    runs anneal.cpp:193-221: at each T_sample grid step, after its energy
    sample, n_sample x (n_sweep sweeps + SSF sample) into that temperature's
    accumulator.  The replayed device schedule equals the same loop issued
    as eager calls (spins and SSF sums bit for bit, energies to rounding:
    the graph takes them from the last colour pass), every T_sample holds
    n_sample samples, and the final spins give each item's own energy."""
    from pyro_cmc_amd import batch

    items = batch.uneven_items(n=5, sizes=(2, 4), seed=8)
    kw = dict(T_hot=5.0, T_cold=0.05, n_steps=8, n_sweep=2, n_burn_in=3, n_sample=2, seed=3)
    Ts = [1.0, 0.2, 0.05]
    res, _ = batch.run_anneal_batch(items, keep_spins=True, T_sample=Ts, **kw)
    grid, s_idx = batch.generate_T_profile(5.0, 0.05, Ts, 8)
    assert len(grid) == 8 + 3 and len(s_idx) == 3  # T_sample merged into the 8-step grid
    for it, r in zip(items, res):
        assert np.array_equal(r.T_sample, Ts) and np.array_equal(r.n_samples, [2, 2, 2])
        assert r.S.shape == (3, 3, it.L ** 3, 2) and r.E.shape == (len(grid),)
        o = oracle.Oracle.pyrochlore(it.L, *it.J)
        o.set_field(np.array(it.B))
        o.set_spins(r.spins)
        assert abs(r.E_final - o.energy_per_cell()) <= 1e-5 * 16 * _scale(it.J, it.B)
    # eager replay of the loop for the L = 4 group
    idx = [i for i, it in enumerate(items) if it.L == 4]
    c = gpu.Context.pyrochlore(4, n_replicas=len(idx), precision=gpu.FP32, seed=3 * 1000003 + 4)
    c.define_J1J2J3(1.0, 1.0, 1.0)
    c.setup()
    c.set_replica_couplings(np.array([items[i].J for i in idx], np.float64))
    for r, i in enumerate(idx):
        c.set_field(np.array(items[i].B, np.float64), r)
    c.randomize(3 * 7919 + 4)
    R = c.R
    c.ssf_configure(("xx", "yy", "zz"), n_slots=3 * R, store_error_term=True)

    def step(T, n_sw):
        c.set_temperatures(np.full(R, T))
        c.sweep(gpu.METROPOLIS, 1, n_sw, count=False)

    step(5.0, 3)
    E = []
    for i, T in enumerate(grid):
        step(T, 2)
        E.append(c.energy_per_cell())
        if i in s_idx:
            j = s_idx.index(i)
            for _ in range(2):
                step(T, 2)
                c.ssf_sample(np.arange(R, dtype=np.int32) + j * R)
    E = np.array(E).T
    for r, i in enumerate(idx):
        assert np.array_equal(res[i].spins, c.get_spins(r))
        assert np.allclose(res[i].E, E[r], rtol=0, atol=1e-5 * 16 * 8)
        for j in range(3):
            S, S2, n = c.ssf_read(j * R + r)
            assert n == 2
            assert np.array_equal(res[i].S[j], S) and np.array_equal(res[i].S2[j], S2)
    c.close()
