"""GPU parity: the HIP library against the CPU oracle on the same inputs.

Fixed configurations (random, ferromagnetic, spiral_byhand) give local
fields, energies and structure factors to 1e-12 (FP64) and 1e-5 (FP32),
relative to the natural scale of each quantity.  Sweeps are compared draw for
draw: the oracle's counter mode visits the product's colours in order with
the same Philox counters, so FP64 trajectories agree to rounding and every
accept/reject decision and lifted direction matches.
"""
import numpy as np
import pytest

import fp32_bound
from helpers import byhand_spiral, random_spins

pytestmark = pytest.mark.gpu

TOL = {64: 1e-12, 32: 1e-5}


def make(gpu, oracle, L, J=(1.0, 0.0, 0.0), Delta=1.0, prec=64, R=1, seed=1234, B=None, T_ref=1.0):
    ctx = gpu.Context.pyrochlore(L, n_replicas=R, precision=prec, seed=seed)
    ctx.define_J1J2J3(*J, Delta)
    ctx.setup()
    if B is not None:
        ctx.set_field(np.asarray(B, float))
    ctx.set_T_ref(T_ref)
    os_ = []
    for r in range(R):
        o = oracle.Oracle.pyrochlore(L, *J, Delta)
        if B is not None:
            o.set_field(np.asarray(B, float))
        o.set_T_ref(T_ref)
        os_.append(o)
    return ctx, os_


def field_scale(J, B=None):
    J1, J2, J3 = J
    return 6 * abs(J1) + 12 * abs(J2) + 12 * abs(J3) + (0 if B is None else np.linalg.norm(B))


CONFIGS = [
    ((1.0, 0.0, 0.0), 1.0, None),
    ((1.0, -0.05, 0.05), 1.0, None),
    ((1.0, 0.1, 0.1), 1.0, (0.0, 0.0, 0.2)),
    ((-1.0, 0.2, -0.1), 0.3, (0.1, -0.4, 0.25)),  # XXZ anisotropy + field
]


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("J,Delta,B", CONFIGS)
@pytest.mark.parametrize("state", ["random", "spiral", "ferro"])
def test_fields_and_energy(gpu, oracle, prec, J, Delta, B, state):
    L = 3
    ctx, (o,) = make(gpu, oracle, L, J, Delta, prec, B=B)
    pos = o.positions()
    if state == "random":
        S = random_spins(ctx.N, np.random.default_rng(42))
    elif state == "spiral":
        S = byhand_spiral(pos, [2 * np.pi / (8 * L), 0, 2 * 2 * np.pi / (8 * L)])
    else:
        S = np.tile(np.array([2.0, 1.0, -2.0]) / 3, (ctx.N, 1))
    ctx.set_spins(S)
    o.set_spins(S)
    h = ctx.local_fields()
    want = o.local_fields()
    scale = field_scale(J)
    assert np.max(np.abs(h - want)) <= TOL[prec] * scale
    e = ctx.energy_per_cell()[0]
    e_ref = o.energy_per_cell()
    e_scale = 16 * 0.5 * field_scale(J, B)  # sum of |terms| bound per cell
    assert abs(e - e_ref) <= TOL[prec] * e_scale


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("L,tile", [(8, None), (16, None), (24, None), (16, (8, 4, 4)), (32, (8, 8, 4)),
                                    (16, (4, 4, 8)), (24, (6, 6, 4))])
def test_energy_tiled(gpu, oracle, prec, L, tile):
    """Lattices the tiled sweep takes get their energies from tiles in LDS
    (one workgroup per tile and replica): against the oracle for XXZ
    couplings, all three shells, a field, and per-replica models."""
    J, Delta, B = (1.0, -0.05, 0.05), 0.7, (0.05, -0.1, 0.2)
    R = 3
    ctx, os_ = make(gpu, oracle, L, J, Delta, prec, R=R, B=B)
    if tile is not None:
        ctx.set_tile_shape(tile)
    assert ctx.tile_shape() != (0, 0, 0)
    rng = np.random.default_rng(L)
    for r in range(R):
        S = random_spins(ctx.N, rng)
        ctx.set_spins(S, r)
        os_[r].set_spins(S)
    e_scale = 16 * 0.5 * field_scale(J, B)
    e = ctx.energy_per_cell()
    for r in range(R):
        assert abs(e[r] - os_[r].energy_per_cell()) <= TOL[prec] * e_scale
    # per-replica models: each replica against an oracle of its own couplings
    Jrep = np.array([[1.0, 0.1, -0.1, 1.0], [-0.5, 0.2, 0.0, 1.3], [1.0, 0.0, 0.15, 0.8]])
    ctx.set_replica_couplings(Jrep)
    e = ctx.energy_per_cell()
    for r in range(R):
        o = oracle.Oracle.pyrochlore(L, *Jrep[r])
        o.set_field(np.asarray(B, float))
        o.set_spins(ctx.get_spins(r))
        assert abs(e[r] - o.energy_per_cell()) <= TOL[prec] * 16 * 0.5 * field_scale(tuple(Jrep[r, :3]), B)


@pytest.mark.parametrize("prec", [64, 32])
def test_randomize(gpu, oracle, prec):
    ctx, os_ = make(gpu, oracle, 2, prec=prec, R=3)
    ctx.randomize(987654321)
    for r, o in enumerate(os_):
        o.randomize_ctr(987654321, r)
        S = ctx.get_spins(r)
        assert np.max(np.abs(S - o.get_spins())) <= (1e-15 if prec == 64 else 1e-7)
        assert np.allclose(np.linalg.norm(S, axis=1), 1.0, atol=1e-6)
        assert np.all(ctx.get_lifted_dir(r) == 1)


def _oracle_sweeps(o, r, kind, n_or, n_sweeps, T, colours, n_col, seed, pass0):
    acc = 0
    p = pass0
    for _ in range(n_sweeps):
        acc += o.pass_ctr(kind, T, colours, n_col, seed, r, p)
        p += 1
        for _ in range(n_or):
            o.pass_ctr(3, T, colours, n_col, seed, r, p)
            p += 1
    return acc


@pytest.mark.parametrize("strategy", [1, 2, 3, 4, 6])
@pytest.mark.parametrize("kind,n_or", [(0, 1), (0, 2), (1, 1), (2, 0), (1, 0)])
@pytest.mark.parametrize("J,Delta,B", [CONFIGS[1], CONFIGS[3]])
def test_sweep_trajectory_fp64(gpu, oracle, strategy, kind, n_or, J, Delta, B):
    L, R, seed = 2, 3, 0xC0FFEE1234
    T = np.array([0.3, 0.8, 2.5])
    ctx, os_ = make(gpu, oracle, L, J, Delta, 64, R=R, seed=seed, B=B, T_ref=0.7)
    ctx.set_strategy(strategy)
    ctx.set_temperatures(T)
    rng = np.random.default_rng(5)
    dirs = []
    for r in range(R):
        S = random_spins(ctx.N, rng)
        d = rng.choice(np.array([-1, 1], np.int8), ctx.N)
        ctx.set_spins(S, r)
        ctx.set_lifted_dir(d, r)
        os_[r].set_spins(S)
        os_[r].set_dir(d)
    colours = ctx.colouring()
    n_col = ctx.info()["n_colours"]
    pass0 = ctx.pass_count()
    acc = ctx.sweep(kind, n_or, 3)
    for r, o in enumerate(os_):
        want_acc = _oracle_sweeps(o, r, kind, n_or, 3, T[r], colours, n_col, seed, pass0)
        assert int(acc[r]) == want_acc
        assert np.max(np.abs(ctx.get_spins(r) - o.get_spins())) < 1e-12
        assert np.array_equal(ctx.get_lifted_dir(r), o.get_dir())
    assert ctx.pass_count() == pass0 + 3 * (1 + n_or)


@pytest.mark.parametrize("kind,n_or", [(0, 2), (1, 1), (2, 1), (1, 0)])
def test_sweep_fp32_tracks_fp64(gpu, oracle, kind, n_or):
    """FP32 passes against the FP64 oracle pass by pass, each from the GPU's
    own state: every site's deviation is bounded by FP32_SITE_TOL unless a
    near-threshold decision upstream of it in the pass explains it."""
    L, seed, T = 3, 77, 0.6
    J = (1.0, -0.05, 0.05)
    ctx, (o,) = make(gpu, oracle, L, J, 1.0, 32, seed=seed)
    ctx.set_temperatures([T])
    S = random_spins(ctx.N, np.random.default_rng(9))
    ctx.set_spins(S)
    colours = ctx.colouring()
    n_col = ctx.info()["n_colours"]
    adj = fp32_bound.padded_adjacency(o, (*J, 1.0))
    sites = np.arange(ctx.N)
    for sweep in range(3):
        for p in range(1 + n_or):
            o.set_spins(ctx.get_spins())
            o.set_dir(ctx.get_lifted_dir())
            pidx = ctx.pass_count()
            if p == 0:
                ctx.sweep(kind, 0, 1)
            else:
                ctx.overrelax(1)
            info = np.empty((ctx.N, 4))
            o.pass_ctr(kind if p == 0 else 3, T, colours, n_col, seed, 0, pidx, info=info)
            E, taint, seeds = fp32_bound.propagate(sites, colours, info, adj, T, field_scale(J), ctx.N)
            fp32_bound.check(ctx.get_spins(), o.get_spins(), E, taint, seeds, ctx.N, f"kind {kind} sweep {sweep} pass {p}")
            if kind in (1, 2) and p == 0:  # a lifted direction may differ only where a decision may flip
                dd = ctx.get_lifted_dir() != o.get_dir()
                assert np.all(taint[dd])
    assert np.allclose(np.linalg.norm(ctx.get_spins(), axis=1), 1.0, atol=1e-6)


@pytest.mark.parametrize("pair,J,R,prec,fused", [((1, 2), (1.0, 0.1, 0.1), 4, 32, False),
                                                  ((3, 4), (1.0, 0.1, 0.1), 4, 32, False),
                                                  ((3, 4), (1.0, -0.2, 0.0), 4, 32, False),
                                                  ((3, 4), (-1.0, 0.0, 0.0), 4, 32, False),
                                                  ((3, 4), (1.0, -0.05, 0.05), 16, 64, False),
                                                  ((1, 2), (1.0, 0.1, 0.0), 8, 64, False),
                                                  # the fused pass order
                                                  ((3, 4), (1.0, -0.05, 0.05), 8, 32, True),
                                                  ((3, 4), (1.0, 0.1, 0.0), 4, 32, True),
                                                  ((3, 4), (1.0, -0.05, 0.05), 4, 64, True),
                                                  # XCD-resident sweeps (one launch, stages
                                                  # synchronised in each XCD's L2)
                                                  ((3, 6), (1.0, 0.1, 0.1), 4, 32, False),
                                                  ((3, 6), (1.0, -0.2, 0.0), 12, 32, False),
                                                  ((3, 6), (-1.0, 0.0, 0.0), 4, 32, False),
                                                  ((3, 6), (1.0, -0.05, 0.05), 16, 64, False),
                                                  ((3, 6), (1.0, -0.05, 0.05), 24, 32, True),
                                                  ((3, 6), (1.0, 0.1, 0.0), 9, 64, True)])
def test_strategies_bit_identical(gpu, pair, J, R, prec, fused):
    """Per-colour launches (1, 3), resident workgroups (2, 4) and the
    XCD-resident launch (6): spins, lifted directions and accept counts of
    both pass mixes agree bit for bit, pass by pass and in the fused pass
    order."""
    res = []
    T = np.geomspace(0.1, 3.0, R)
    for strategy in pair:
        ctx = gpu.Context.pyrochlore(3, n_replicas=R, precision=prec, seed=99)
        ctx.define_J1J2J3(*J)
        ctx.setup()
        ctx.set_fused_passes(fused)
        ctx.set_strategy(strategy)
        ctx.set_temperatures(T)
        ctx.randomize(5)
        acc = ctx.sweep(gpu.LIFTED, 1, 4)
        acc2 = ctx.sweep(gpu.METROPOLIS, 2, 3)
        res.append((acc, acc2, [ctx.get_spins(r) for r in range(R)], [ctx.get_lifted_dir(r) for r in range(R)]))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])
    for a, b in zip(res[0][2], res[1][2]):
        assert np.array_equal(a, b)
    for a, b in zip(res[0][3], res[1][3]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", [64, 32])
def test_stencil_and_table_kernels_agree(gpu, prec):
    """The compile-time-stencil kernels sum the field by shell, the table
    kernels slot by slot: same visit order and draws, equal to rounding."""
    out = []
    for strategy in (1, 3):
        ctx = gpu.Context.pyrochlore(3, n_replicas=2, precision=prec, seed=5)
        ctx.define_J1J2J3(1.0, -0.05, 0.05, 0.7)
        ctx.setup()
        ctx.set_strategy(strategy)
        ctx.set_temperatures([0.4, 1.5])
        ctx.randomize(8)
        ctx.sweep(gpu.METROPOLIS, 1, 1)
        out.append(ctx.get_spins(1))
    tol = 1e-12 if prec == 64 else 1e-4
    assert np.mean(np.abs(out[0] - out[1]).max(1) > tol) < 0.01


@pytest.mark.parametrize("prec", [32, 64])
def test_stencil_strategies_sample_identical_energies(gpu, prec):
    """Tempering rounds and anneal steps take their energies from the last
    sweep pass under both stencil strategies (per-colour launches, 3, and
    resident workgroups, 4) and under the automatic choice between them:
    spins, sampled E / E^2 sums, swap statistics and anneal energy
    histories agree bit for bit, so the autotuner's timing never reaches a
    result."""
    out = []
    for strategy in (3, 4, 6, 0):
        ctx = gpu.Context.pyrochlore(3, n_replicas=6, precision=prec, seed=31)
        ctx.define_J1J2J3(1.0, -0.05, 0.05, 0.9)
        ctx.setup()
        ctx.set_field(np.array([0.05, 0.0, -0.1]))
        ctx.set_strategy(strategy)
        ctx.randomize(2)
        ctx.pt_init(np.geomspace(2.0, 0.1, 6), 5)
        ctx.pt_rounds(gpu.METROPOLIS, 2, 2, n_rounds=5)
        st = ctx.pt_stats()
        ctx.anneal_schedule(np.geomspace(3.0, 0.2, 4))
        ctx.anneal_steps(gpu.LIFTED, 1, 2, False, 4)
        hist = ctx.anneal_history()
        out.append((st, hist, [ctx.get_spins(r) for r in range(ctx.R)]))
        ctx.close()
    for st, hist, spins in out[1:]:
        for key in ("E", "E2", "n", "attempts", "accepts"):
            assert np.array_equal(st[key], out[0][0][key]), key
        assert np.array_equal(hist, out[0][1])
        for a, b in zip(spins, out[0][2]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("stencil", [True, False])
def test_autotuner_keeps_results(gpu, oracle, stencil):
    """Auto-tuning picks within a bit-identical pair (3/4 for the stencil
    kernels, 1/2 for the table kernels), so results do not depend on it."""
    out = []
    for strategy in (0, 3 if stencil else 1):
        ctx = gpu.Context.pyrochlore(2, n_replicas=8, precision=32, seed=3)
        if stencil:
            ctx.define_J1J2J3(1.0)
        else:
            for name, rel, J in oracle.j1j2j3_couplings(1.0, 0.0, 0.0):
                ctx.define_coupling(name, rel, J)
        ctx.setup()
        ctx.set_strategy(strategy)
        ctx.set_temperatures(np.linspace(0.2, 2, 8))
        ctx.randomize(1)
        acc = ctx.sweep(gpu.METROPOLIS, 2, 3)
        out.append((acc, ctx.get_spins(7), ctx.pass_count()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


@pytest.mark.parametrize("prec", [32, 64])
def test_autotuner_keeps_energy_sums(gpu, prec):
    """Tempering rounds and anneal steps take their energies from the last
    colour pass under strategies 3 and 4 alike: the same partial per wave of
    64 cells and sublattice, summed by the same finish kernel, so E, E^2,
    the swaps and the anneal history are bit-identical whichever of the
    pair the tuner picks (L = 12: 27 partials per sublattice)."""
    out = []
    for strategy in (0, 3, 4):
        ctx = gpu.Context.pyrochlore(12, n_replicas=8, precision=prec, seed=5)
        ctx.define_J1J2J3(1.0, -0.05, 0.05)
        ctx.setup()
        ctx.set_strategy(strategy)
        ctx.randomize(2)
        ctx.pt_init(np.geomspace(2.0, 0.2, 8), 3)
        ctx.pt_rounds(gpu.METROPOLIS, 2, 2, False, 6)
        st = ctx.pt_stats()
        ctx.anneal_schedule(np.geomspace(1.0, 0.1, 5))
        ctx.anneal_steps(gpu.LIFTED, 1, 2, False, 5)
        out.append((st["E"], st["E2"], st["accepts"], ctx.pt_permutation(), ctx.anneal_history(),
                    ctx.get_spins(7)))
        ctx.close()
    for o in out[1:]:
        for a, b in zip(out[0], o):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", [64, 32])
def test_tempering_rounds(gpu, oracle, prec):
    L, R, swap_seed = 2, 6, 4242
    T = np.geomspace(3.0, 0.2, R)
    ctx, os_ = make(gpu, oracle, L, (1.0, 0.1, 0.0), 1.0, prec, R=R)
    rng = np.random.default_rng(1)
    for r in range(R):
        S = random_spins(ctx.N, rng) if r % 2 else np.tile([0, 0, 1.0], (ctx.N, 1))
        ctx.set_spins(S, r)
        os_[r].set_spins(S)
    ctx.pt_init(T, swap_seed)
    e = np.array([o.energy_per_cell() for o in os_])
    rep = np.arange(R, dtype=np.int32)
    att = np.zeros(R, np.uint64)
    accp = np.zeros(R, np.uint64)
    Es = np.zeros(R)
    E2s = np.zeros(R)
    n_rounds = 7
    for rnd in range(n_rounds):
        for s in range(R):
            Es[s] += e[rep[s]]
            E2s[s] += e[rep[s]] ** 2
        rep = oracle.exchange(T, e, L ** 3, rep, rnd, swap_seed, att, accp)
    ctx.pt_rounds(n_sweep=0, n_rounds=n_rounds)
    st = ctx.pt_stats()
    assert np.array_equal(ctx.pt_permutation(), rep)
    assert np.array_equal(st["attempts"], att[: R - 1])
    assert np.array_equal(st["accepts"], accp[: R - 1])
    assert np.all(st["n"] == n_rounds)
    tol = TOL[prec] * 16 * 8
    assert np.allclose(st["E"], Es, rtol=0, atol=tol * n_rounds)
    assert np.allclose(st["E2"], E2s, rtol=0, atol=tol * n_rounds * 100)


def test_tempering_exchange_interval(gpu, oracle):
    """With an exchange every 3rd round, exchange x (at the end of rounds 2,
    5, 8, ...) has parity x mod 2 and counter x; the other rounds only
    sample energies.  Spins stay fixed (no sweeps), so the oracle replays
    the permutation and statistics draw for draw."""
    L, R, swap_seed, every = 2, 6, 4242, 3
    T = np.geomspace(3.0, 0.2, R)
    ctx, os_ = make(gpu, oracle, L, (1.0, 0.1, 0.0), 1.0, 64, R=R)
    rng = np.random.default_rng(5)
    for r in range(R):
        S = random_spins(ctx.N, rng)
        ctx.set_spins(S, r)
        os_[r].set_spins(S)
    ctx.pt_init(T, swap_seed)
    ctx.pt_exchange_interval(every)
    e = np.array([o.energy_per_cell() for o in os_])
    rep = np.arange(R, dtype=np.int32)
    att = np.zeros(R, np.uint64)
    accp = np.zeros(R, np.uint64)
    n_rounds = 10
    for rnd in range(n_rounds):
        if (rnd + 1) % every == 0:
            rep = oracle.exchange(T, e, L ** 3, rep, rnd // every, swap_seed, att, accp)
    ctx.pt_rounds(n_sweep=0, n_rounds=n_rounds)
    st = ctx.pt_stats()
    assert np.array_equal(ctx.pt_permutation(), rep)
    assert np.array_equal(st["attempts"], att[: R - 1])
    assert np.array_equal(st["accepts"], accp[: R - 1])
    assert att.sum() > 0 and accp.sum() > 0  # the test must see swaps
    assert np.all(st["n"] == n_rounds)  # energies are sampled every round
    with pytest.raises(gpu.PyroError):
        ctx.pt_exchange_interval(0)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize(
    "L,R,J,kind,n_or",
    [
        (3, 5, (1.0, -0.05, 0.05), 0, 2),  # 16 colours; 27 cells: one partly filled block
        (3, 4, (1.0, 0.1, 0.0), 1, 1),  # 4 colours, lifted
        (4, 3, (1.0, 0.1, 0.1), 0, 0),  # no OR pass: the Metropolis pass is the last
        (12, 64, (1.0, -0.05, 0.05), 0, 2),  # the L=12 x 64 tempering lattice, 27 blocks
        # the fused pass order: the energy pass is OR pass 2 on its own, or
        # with one OR pass the launch of the first kind and the OR update
        (12, 64, (1.0, -0.05, 0.05), 0, -2),
        (3, 4, (1.0, 0.1, 0.0), 1, -1),
    ],
)
@pytest.mark.parametrize("strategy", [3, 6])
def test_tempering_fused_energy(gpu, oracle, prec, L, R, J, kind, n_or, strategy):
    """With per-colour stencil launches the round's last colour pass also
    leaves the energy partials (bonds to earlier colours, counted when the
    later endpoint is updated), replacing the energy launch.  The energies
    each round samples equal those of the state the round leaves: from the
    energy kernel and from the oracle on the downloaded spins."""
    T = np.geomspace(2.0, 0.05, R)
    ctx, os_ = make(gpu, oracle, L, J, 0.8, prec, R=R, B=(0.1, -0.2, 0.3))
    ctx.set_strategy(strategy)
    if n_or < 0:  # negative: that many OR passes in the fused pass order
        n_or = -n_or
        ctx.set_fused_passes(True)
    ctx.randomize(11)
    ctx.pt_init(T, 9)
    try:
        ctx.pt_rounds(kind, n_or, 2, n_rounds=1)
    except gpu.PyroError as e:
        assert strategy == 6 and e.code == 5  # the XCD-resident grid does not fit at once here
        pytest.skip(str(e))
    st = ctx.pt_stats()
    assert np.all(st["n"] == 1)
    e_dev = ctx.energy_per_cell()
    e_orc = []
    for r in range(R):
        o = os_[0]
        o.set_spins(ctx.get_spins(r))
        e_orc.append(o.energy_per_cell())
    # the exchange permutes slots, not spins: compare the sampled multiset
    tol = TOL[prec] * 16 * 8
    assert np.allclose(np.sort(st["E"]), np.sort(e_dev), rtol=0, atol=tol)
    assert np.allclose(np.sort(st["E"]), np.sort(e_orc), rtol=0, atol=tol)
    assert np.ptp(st["E"]) > 100 * tol  # distinct energies, so the sort pairs them
    # further rounds (the second replays the captured round): each adds the
    # energies of the state it leaves, whatever the slot order
    prev = st["E"].sum()
    for n in (2, 3):
        ctx.pt_rounds(kind, n_or, 1, n_rounds=1)
        s = ctx.pt_stats()
        assert np.all(s["n"] == n)
        assert abs(s["E"].sum() - prev - ctx.energy_per_cell().sum()) < tol * R
        prev = s["E"].sum()


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("fused,mapped", [(True, False), (False, True), (True, True)])
def test_ssf_rounds_equal_across_strategies(gpu, oracle, prec, fused, mapped):
    """Tempering rounds with SSF samples, over calls of 1, 2 and 5 rounds
    (the captured round replayed across calls): the XCD-resident sweep and per-colour launches sweep bit-identically, so
    the SSF sums, the energy and swap statistics and the spins are
    identical, with and without a slot map (one that reverses the ladder
    into the upper accumulators and skips the hottest slot)."""
    L, R = 6, 16
    T = np.geomspace(2.0, 0.05, R)
    n_slots = 2 * R if mapped else R
    out = []
    for strategy in (6, 3):
        ctx, _ = make(gpu, oracle, L, (1.0, -0.05, 0.05), 0.8, prec, R=R, B=(0.1, -0.2, 0.3))
        ctx.set_strategy(strategy)
        ctx.set_fused_passes(fused)
        ctx.randomize(21)
        ctx.pt_init(T, 5)
        ctx.ssf_configure(("xx", "zz", "xy"), n_slots=n_slots, store_error_term=True)
        if mapped:
            m = (2 * R - 1 - np.arange(R)).astype(np.int32)
            m[0] = -1
            ctx.pt_ssf_map(m)
        for n in (1, 2, 5):
            ctx.pt_rounds(gpu.METROPOLIS, 2, 2, True, n)
        out.append((ctx.pt_stats(), [ctx.ssf_read(s) for s in range(n_slots)],
                    [ctx.get_spins(r) for r in range(R)]))
        ctx.close()
    (sa, qa, pa), (sb, qb, pb) = out
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    for x, y in zip(qa, qb):
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) and x[2] == y[2]
    assert sum(q[2] for q in qa) == 8 * (R - 1 if mapped else R)  # every round sampled
    for x, y in zip(pa, pb):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("kind,n_or", [(0, 2), (1, 0), (2, 1), (1, 1), (0, 0), (2, 0)])
@pytest.mark.parametrize("per_rep", [True, False])
@pytest.mark.parametrize("strategy", [3, 4, 6, "3-fused", "4-fused", "6-fused"])
def test_energy_pass_is_bit_identical(gpu, prec, kind, n_or, per_rep, strategy):
    """The kernels that also sum the energy (the last pass of an anneal step
    or tempering round; per-colour launches and resident workgroups) update
    every site exactly as the plain kernels do: anneal steps and plain
    sweeps end in the same spins, bit for bit, for every update kind as the
    last pass, with XXZ per-replica couplings and fields or the single-model
    kernels."""
    out = []
    for fused in (True, False):
        ctx = gpu.Context.pyrochlore(3, n_replicas=3, precision=prec, seed=21)
        if per_rep:
            ctx.define_J1J2J3(1.0, 0.1, 0.1)
        else:
            ctx.define_J1J2J3(1.0, -0.05, 0.05, 0.8)
        ctx.setup()
        if per_rep:
            ctx.set_replica_couplings(np.array([[1.0, -0.2, 0.15, 0.7], [-1.0, 0.1, -0.05, 1.3],
                                                [0.5, 0.2, 0.1, 1.0]]))
            for r, B in enumerate([(0.0, 0.0, 0.3), (0.1, -0.3, 0.2), (0.0, 0.0, 0.0)]):
                ctx.set_field(np.array(B), r)
        else:
            ctx.set_field(np.array([0.1, -0.2, 0.3]))
        if isinstance(strategy, str):  # the fused pass order's energy pass
            ctx.set_strategy(int(strategy[0]))
            ctx.set_fused_passes(True)
        else:
            ctx.set_strategy(strategy)
        ctx.randomize(8)
        if fused:
            ctx.anneal_schedule(np.array([1.5, 0.6]))
            ctx.anneal_steps(kind, n_or, 2, False, 2)
        else:
            for T in (1.5, 0.6):
                ctx.set_temperatures(np.full(ctx.R, T))
                ctx.sweep(kind, n_or, 2, count=False)
        out.append([(ctx.get_spins(r), ctx.get_lifted_dir(r)) for r in range(ctx.R)])
    for (Sa, da), (Sb, db) in zip(*out):
        assert np.array_equal(Sa, Sb)
        assert np.array_equal(da, db)


def test_tempering_ssf_slot_map(gpu, oracle):
    """SSF samples taken inside tempering rounds land in the accumulator the
    map names for the rung a replica occupies at that round (temper.cpp:426-429:
    only the T_sample rungs sample)."""
    L, R, swap_seed = 2, 5, 77
    T = np.geomspace(2.0, 0.1, R)
    ctx, os_ = make(gpu, oracle, L, (1.0, 0.1, 0.0), 1.0, 64, R=R)
    rng = np.random.default_rng(4)
    for r in range(R):
        S = random_spins(ctx.N, rng)
        ctx.set_spins(S, r)
        os_[r].set_spins(S)
    corr = ("xx", "yz")
    ssf = [o.ssf_sample(corr) for o in os_]
    e = np.array([o.energy_per_cell() for o in os_])
    ctx.pt_init(T, swap_seed)
    ctx.ssf_configure(corr, n_slots=2)
    mapping = np.array([-1, 1, -1, 0, -1], np.int32)
    ctx.pt_ssf_map(mapping)
    rep = np.arange(R, dtype=np.int32)
    att = np.zeros(R, np.uint64)
    accp = np.zeros(R, np.uint64)
    want = np.zeros((2, len(corr), L ** 3, 2))
    n_rounds = 6
    for rnd in range(n_rounds):
        for s in range(R):
            if mapping[s] >= 0:
                want[mapping[s]] += ssf[rep[s]]
        rep = oracle.exchange(T, e, L ** 3, rep, rnd, swap_seed, att, accp)
    ctx.pt_rounds(n_sweep=0, sample_ssf=True, n_rounds=n_rounds)
    assert np.array_equal(ctx.pt_permutation(), rep)
    assert accp.sum() > 0  # the test must see swaps
    for j in range(2):
        S, _, n = ctx.ssf_read(j)
        assert n == n_rounds
        assert np.max(np.abs(S - want[j])) <= 1e-12 * ctx.N ** 2 * n_rounds
    ctx.pt_ssf_map(None)
    with pytest.raises(gpu.PyroError):
        ctx.pt_rounds(n_sweep=0, sample_ssf=True, n_rounds=1)  # 2 slots < 5 rungs, no map


def test_tempering_graph_equals_eager(gpu):
    out = []
    for prof in (False, True):
        ctx = gpu.Context.pyrochlore(2, n_replicas=8, precision=32, seed=17)
        ctx.define_J1J2J3(1.0, -0.05, 0.05)
        ctx.setup()
        ctx.randomize(3)
        ctx.pt_init(np.geomspace(2.0, 0.05, 8), 9)
        ctx.ssf_configure(("xx", "yy", "zz"), n_slots=8)
        ctx.profile(prof)
        ctx.pt_rounds(gpu.METROPOLIS, 2, 4, True, 5)
        st = ctx.pt_stats()
        out.append((ctx.pt_permutation(), st["E"], ctx.get_spins(3), ctx.ssf_read(2)[0]))
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", [64, 32])
def test_ssf_contracted(gpu, oracle, prec):
    L, R = 3, 2
    corr = ("xx", "yy", "zz", "xy", "xz", "yz", "yx", "zx", "zy")
    ctx, os_ = make(gpu, oracle, L, prec=prec, R=R)
    ctx.ssf_configure(corr, n_slots=2, store_error_term=True)
    rng = np.random.default_rng(8)
    want = np.zeros((2, len(corr), L ** 3, 2))
    want2 = np.zeros_like(want)
    pos = os_[0].positions()
    for step in range(3):
        for r in range(R):
            S = random_spins(ctx.N, rng) if step < 2 else byhand_spiral(pos, [0, 0, 2 * np.pi / (8 * L)])
            ctx.set_spins(S, r)
            os_[r].set_spins(S)
            s = os_[r].ssf_sample(corr)
            want[1 - r] += s
            want2[1 - r] += s ** 2
        ctx.ssf_sample([1, 0])  # replica 0 -> slot 1, replica 1 -> slot 0
    N2 = ctx.N ** 2
    for slot in range(2):
        S, S2, n = ctx.ssf_read(slot)
        assert n == 3
        assert np.max(np.abs(S - want[slot])) <= TOL[prec] * N2 * 3
        assert np.max(np.abs(S2 - want2[slot])) <= TOL[prec] * N2 * N2 * 3 * 4
    ft = ctx.spin_ft(1)
    assert np.max(np.abs(ft - os_[1].spin_ft())) <= TOL[prec] * L ** 3


@pytest.mark.parametrize("prec", [64, 32])
def test_ssf_contracted_beyond_sixteen_sublattices(gpu, prec):
    """A cell of 20 sublattices: more than the 16 whose values the
    contraction loads together (the rest go one by one), and in FP32 more
    than the staged kernel's 3 x 16 features, so the generic contraction
    runs.  The correlators against conj(b^a) b^b with
    b^a(K) = sum_mu e^{-i q(K).r_mu} A^a_mu(K), from the per-sublattice
    transform A, q(K) = 2 pi K_c / (cell L) for the centred index K_c."""
    n_sl, cell, L = 20, np.array([4, 4, 4]), np.array([3, 4, 5])
    pos = np.array([(i % 4, (i // 4) % 4, i // 16) for i in range(n_sl)])
    ctx = gpu.Context(pos, [0] * n_sl, cell, L, n_replicas=1, precision=prec, seed=5)
    ctx.define_coupling("J1", [[(4, 0, 0), (-4, 0, 0)]], np.eye(3))
    ctx.setup()
    corr = ("xx", "xy", "zy", "yz")
    ctx.ssf_configure(corr, n_slots=1, store_error_term=False)
    ctx.set_spins(random_spins(ctx.N, np.random.default_rng(20)), 0)
    ctx.ssf_sample([0])
    S, _, n = ctx.ssf_read(0, with_sq=False)
    assert n == 1
    ft = ctx.spin_ft(0)
    A = ft[..., 0] + 1j * ft[..., 1]  # [a][mu][K]
    k = np.stack(np.meshgrid(*[np.arange(l) for l in L], indexing="ij"), -1).reshape(-1, 3)
    kc = np.where(2 * k < L, k, k - L)
    arg = 2 * np.pi * (kc / (cell * L)) @ pos.T  # [K][mu]
    b = np.einsum("km,amk->ak", np.exp(-1j * arg), A)
    ax = {"x": 0, "y": 1, "z": 2}
    for c, (al, be) in enumerate(corr):
        want = np.conj(b[ax[al]]) * b[ax[be]]
        assert np.max(np.abs(S[c, :, 0] - want.real)) <= TOL[prec] * ctx.N ** 2 * 3
        assert np.max(np.abs(S[c, :, 1] - want.imag)) <= TOL[prec] * ctx.N ** 2 * 3


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("L", [4, 5, 8, 12, 16])
def test_spin_ft_against_direct_dft(gpu, oracle, prec, L):
    """Per-sublattice DFT (save_ft_spin_state layout): through the per-axis
    path at line lengths that pad the FP32 GEMM tiles (5), and through the
    fused whole-field kernel on the cubic edges it is built for (4, 8, 12,
    16; 12 is the tempering lattice)."""
    ctx, os_ = make(gpu, oracle, L, prec=prec)
    S = random_spins(ctx.N, np.random.default_rng(L))
    ctx.set_spins(S, 0)
    os_[0].set_spins(S)
    ft = ctx.spin_ft(0)
    want = os_[0].spin_ft()
    assert np.max(np.abs(ft - want)) <= TOL[prec] * L ** 3


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("L", [(32, 3, 2), (2, 48, 3), (3, 2, 64), (48, 2, 4), (2, 3, 32), (4, 64, 2), (2, 2, 48)])
def test_spin_ft_radix_lines(gpu, oracle, prec, L):
    """Line lengths 32, 48 and 64 on each axis (strided and contiguous lines):
    the FP32 Cooley-Tukey kernel (16-point GEMMs, then twiddles and 2-, 3- or
    4-point butterflies) against the direct DFT."""
    ctx, os_ = make(gpu, oracle, L, prec=prec)
    S = random_spins(ctx.N, np.random.default_rng(sum(L)))
    ctx.set_spins(S, 0)
    os_[0].set_spins(S)
    ft = ctx.spin_ft(0)
    want = os_[0].spin_ft()
    assert np.max(np.abs(ft - want)) <= TOL[prec] * np.prod(L)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("L", [(2, 32, 48), (3, 48, 32), (2, 64, 64), (4, 48, 48), (48, 48, 48)])
def test_spin_ft_planes(gpu, oracle, prec, L):
    """L1 and L2 of 32, 48 or 64 points: axes 2 and 1 of each plane in one FP32
    kernel, then axis 0.  The direct DFT of the oracle is too slow at these
    sizes, so the check is numpy's FFT of each (component, sublattice) field;
    the same FFT is first pinned to the oracle's direct DFT (sign, order,
    no normalisation) on a small lattice."""
    small = (2, 3, 4)
    ctx, os_ = make(gpu, oracle, small, prec=64)
    S = random_spins(ctx.N, np.random.default_rng(1))
    os_[0].set_spins(S)
    fft = np.fft.fftn(S.T.reshape(3, 16, *small), axes=(2, 3, 4)).reshape(3, 16, -1)
    assert np.max(np.abs(fft.real - os_[0].spin_ft()[..., 0])) < 1e-12
    assert np.max(np.abs(fft.imag - os_[0].spin_ft()[..., 1])) < 1e-12
    ctx.close()

    ctx = gpu.Context.pyrochlore(L, precision=prec, seed=3)
    ctx.define_J1J2J3(1.0)
    ctx.setup()
    S = random_spins(ctx.N, np.random.default_rng(sum(L)))
    ctx.set_spins(S, 0)
    ft = ctx.spin_ft(0)
    want = np.fft.fftn(S.T.reshape(3, 16, *L), axes=(2, 3, 4)).reshape(3, 16, -1)
    nc = int(np.prod(L))
    assert np.max(np.abs(ft[..., 0] - want.real)) <= TOL[prec] * nc
    assert np.max(np.abs(ft[..., 1] - want.imag)) <= TOL[prec] * nc


@pytest.mark.parametrize("prec", [64, 32])
def test_spin_ft_parseval_full_size(gpu, prec):
    """L = 48 (the large tempering lattice): sum_K |A_mu(K)|^2 = L^3 sum_n |S_mu(n)|^2
    per component and sublattice, and A(0) = sum_n S(n)."""
    L = 48
    ctx = gpu.Context.pyrochlore(L, precision=prec, seed=3)
    ctx.define_J1J2J3(1.0)
    ctx.setup()
    ctx.randomize(5)
    S = ctx.get_spins(0)  # [N][3], site = sl * L^3 + cell
    ft = ctx.spin_ft(0)   # [3][16][L^3][2]
    Sf = S.T.reshape(3, 16, L ** 3)
    power = (ft ** 2).sum(axis=(2, 3))
    want = L ** 3 * (Sf ** 2).sum(axis=2)
    tol = 1e-12 if prec == 64 else 2e-5
    assert np.max(np.abs(power / want - 1)) <= tol
    assert np.max(np.abs(ft[:, :, 0, 0] - Sf.sum(axis=2))) <= tol * L ** 3
    assert np.max(np.abs(ft[:, :, 0, 1])) <= tol * L ** 3


def test_ssf_shared_slot_layers(gpu, oracle):
    """Several replicas feeding one slot are accumulated without races."""
    L, R = 2, 4
    ctx, os_ = make(gpu, oracle, L, R=R)
    ctx.ssf_configure(("xx", "zz"), n_slots=1)
    rng = np.random.default_rng(2)
    want = 0
    for r in range(R):
        S = random_spins(ctx.N, rng)
        ctx.set_spins(S, r)
        os_[r].set_spins(S)
        want = want + os_[r].ssf_sample(("xx", "zz"))
    ctx.ssf_sample([0, 0, 0, 0])
    S, _, n = ctx.ssf_read(0)
    assert n == R
    assert np.max(np.abs(S - want)) <= 1e-12 * ctx.N ** 2 * R


@pytest.mark.parametrize("prec", [64, 32])
def test_ssf_munu_mfma(gpu, oracle, prec):
    L, R = 2, 2
    corr = ("xx", "yz", "zz", "zx")  # zx: a lower tile, read back by conjugation
    ctx, os_ = make(gpu, oracle, L, prec=prec, R=R)
    ctx.ssf_configure(corr, n_slots=2, munu=True, munu_buffer=3)
    rng = np.random.default_rng(11)
    want = np.zeros((2, len(corr), L ** 3, 16, 16, 2))
    n_samp = 5  # > buffer: one flush in the middle, one at read
    for _ in range(n_samp):
        for r in range(R):
            S = random_spins(ctx.N, rng)
            ctx.set_spins(S, r)
            os_[r].set_spins(S)
            _, m = os_[r].ssf_sample(corr, munu=True)
            want[r] += m
        ctx.ssf_sample([0, 1])
    # FP32 products: error ~ 1e-7 relative to |A|^2 ~ L^3 per term
    tol = 2e-5 * L ** 3 * n_samp
    for slot in range(2):
        for c in range(len(corr)):
            got = ctx.ssf_read_munu(slot, c)
            assert np.max(np.abs(got - want[slot, c])) <= tol


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("L,J,kind,n_or,tile", [(8, (1.0, -0.05, 0.05), 0, 2, None), (8, (1.0, 0.1, 0.0), 1, 1, None),
                                                (24, (1.0, 0.1, 0.1), 0, 1, None), (8, (-1.0, 0.2, -0.1), 2, 1, None),
                                                (16, (1.0, 0.1, 0.1), 1, 1, (8, 4, 4)),
                                                (16, (1.0, -0.05, 0.05), 0, 2, (4, 4, 8)),
                                                (16, (1.0, 0.1, 0.0), 2, 1, (8, 8, 4)),
                                                (24, (1.0, 0.1, 0.1), 1, 1, (4, 4, 4)),
                                                # the large tempering run's sweep: lifted + OR on 6x6x4 tiles
                                                (24, (1.0, 0.1, 0.1), 1, 1, (6, 6, 4)),
                                                # a site's passes back to back (pyro_set_fused_passes)
                                                (24, (1.0, 0.1, 0.1), 1, 1, "fused"),
                                                (8, (1.0, -0.05, 0.05), 0, 2, "fused"),
                                                (16, (1.0, 0.1, 0.0), 2, 1, "fused")])
def test_tiled_sweep_against_oracle_order(gpu, oracle, prec, L, J, kind, n_or, tile):
    """The tiled schedule (tiles in LDS, all passes per tile) equals the
    sequential order pyro_sweep_order reports, for every tile shape and
    with fused passes; the oracle replays it draw for draw."""
    R, seed = 2, 0xBEEF
    T = [0.35, 1.7]
    ctx, os_ = make(gpu, oracle, L, J, 1.0, prec, R=R, seed=seed, B=(0.05, -0.1, 0.2))
    fused = tile == "fused"
    if fused:
        ctx.set_fused_passes(True)
        tile = None
    if tile is not None:
        try:
            ctx.set_tile_shape(tile)
        except gpu.PyroError as e:
            assert e.code == 5  # tile + halo exceeds this device's LDS limit
            pytest.skip(str(e))
        assert ctx.tile_shape() == tile
    ctx.set_strategy(5)
    ctx.set_temperatures(T)
    rng = np.random.default_rng(3)
    for r in range(R):
        S = random_spins(ctx.N, rng)
        ctx.set_spins(S, r)
        os_[r].set_spins(S)
    site, koff = ctx.sweep_order(n_or)
    assert np.array_equal(np.sort(site[koff == 0]), np.arange(ctx.N))
    if prec == 64:
        try:
            acc = ctx.sweep(kind, n_or, 2)
        except gpu.PyroError as e:
            assert e.code == 5  # tile + halo exceeds this device's LDS limit
            pytest.skip(str(e))
        for r, o in enumerate(os_):
            a = o.sweep_order(kind, T[r], site, koff, seed, r, 0, shared=fused)
            a += o.sweep_order(kind, T[r], site, koff, seed, r, 1 + n_or, shared=fused)
            diff = np.abs(ctx.get_spins(r) - o.get_spins()).max(1)
            # every decision equal; rounding of the shell-grouped field sum
            # against the oracle's per-bond sum accumulates over the passes
            # (most at lifted moves close to the field axis, where the
            # displacement divides by 1 - u^2; the worst of 27,648 sites at
            # L = 24 reaches 1.1e-10)
            assert int(acc[r]) == a
            assert diff.max() < 1e-9
        return
    # FP32: sweep by sweep from the GPU's own state, every site within its
    # first-order error bound
    adj = fp32_bound.padded_adjacency(os_[0], (*J, 1.0))
    stage = fp32_bound.tiled_stages(ctx, site, koff, fused)
    for sw in range(2):
        before = [(ctx.get_spins(r), ctx.get_lifted_dir(r)) for r in range(R)]
        p0 = ctx.pass_count()
        try:
            ctx.sweep(kind, n_or, 1)
        except gpu.PyroError as e:
            assert e.code == 5  # tile + halo exceeds this device's LDS limit
            pytest.skip(str(e))
        for r, o in enumerate(os_):
            o.set_spins(before[r][0])
            o.set_dir(before[r][1])
            info = np.empty((len(site), 4))
            o.sweep_order(kind, T[r], site, koff, seed, r, p0, info=info, shared=fused)
            E, taint, seeds = fp32_bound.propagate(site, stage, info, adj, T[r], field_scale(J, (0.05, -0.1, 0.2)),
                                                   ctx.N)
            fp32_bound.check(ctx.get_spins(r), o.get_spins(), E, taint, seeds, len(site),
                             f"L {L} tile {ctx.tile_shape()} replica {r}")


@pytest.mark.parametrize("strategy", [3, 4, 6])
@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("kind,n_or", [(0, 1), (0, 2), (1, 1), (2, 1), (1, 2), (0, 3), (1, 0)])
@pytest.mark.parametrize("J,Delta,B", [CONFIGS[1], CONFIGS[3]])
def test_colour_fused_order_against_oracle(gpu, oracle, strategy, prec, kind, n_or, J, Delta, B):
    """Per-colour launches (strategy 3) and resident workgroups (strategy 4)
    in the fused pass order: per colour each site's first-kind and first OR
    update in one stage from one field, then OR passes 2..n_or on their own,
    the sweep's uniforms drawn in order from its blocks.  pyro_sweep_order reports that order, and the oracle
    replays it draw for draw: FP64 decisions, accept counts and lifted
    directions exact, FP32 every site within its first-order error bound."""
    L, R, seed = 3, 3, 0x5EED1234
    T = np.array([0.3, 0.8, 2.5])
    ctx, os_ = make(gpu, oracle, L, J, Delta, prec, R=R, seed=seed, B=B, T_ref=0.7)
    ctx.set_strategy(strategy)
    ctx.set_fused_passes(True)
    ctx.set_temperatures(T)
    rng = np.random.default_rng(17)
    for r in range(R):
        S = random_spins(ctx.N, rng)
        d = rng.choice(np.array([-1, 1], np.int8), ctx.N)
        ctx.set_spins(S, r)
        ctx.set_lifted_dir(d, r)
        os_[r].set_spins(S)
        os_[r].set_dir(d)
    site, koff = ctx.sweep_order(n_or)
    assert len(site) == ctx.N * (1 + n_or)
    for k in range(1 + n_or):
        assert np.array_equal(np.sort(site[koff == k]), np.arange(ctx.N))
    if n_or >= 1:  # each site's first OR update right after its first-kind update
        first = np.flatnonzero(koff == 0)
        assert np.all(koff[first + 1] == 1) and np.array_equal(site[first + 1], site[first])
    p0 = ctx.pass_count()
    if prec == 64:
        acc = ctx.sweep(kind, n_or, 3)
        assert ctx.pass_count() == p0 + 3 * (1 + n_or)
        for r, o in enumerate(os_):
            a = sum(o.sweep_order(kind, T[r], site, koff, seed, r, p0 + sw * (1 + n_or), shared=True)
                    for sw in range(3))
            assert int(acc[r]) == a
            # every decision equal; rounding of the shell-grouped field sum
            # against the oracle's per-bond sum grows over the 3 sweeps of up
            # to 4 passes (most at lifted moves close to the field axis, where
            # the displacement divides by 1 - u^2), as on the tiled order
            assert np.max(np.abs(ctx.get_spins(r) - o.get_spins())) < 1e-10
            assert np.array_equal(ctx.get_lifted_dir(r), o.get_dir())
        return
    adj = fp32_bound.padded_adjacency(os_[0], (*J, Delta))
    stage = fp32_bound.colour_stages(ctx, site, koff, True)
    for sw in range(3):
        before = [(ctx.get_spins(r), ctx.get_lifted_dir(r)) for r in range(R)]
        p0 = ctx.pass_count()
        ctx.sweep(kind, n_or, 1)
        for r, o in enumerate(os_):
            o.set_spins(before[r][0])
            o.set_dir(before[r][1])
            info = np.empty((len(site), 4))
            o.sweep_order(kind, T[r], site, koff, seed, r, p0, info=info, shared=True)
            E, taint, seeds = fp32_bound.propagate(site, stage, info, adj, T[r], field_scale(J, B), ctx.N)
            fp32_bound.check(ctx.get_spins(r), o.get_spins(), E, taint, seeds, len(site),
                             f"fused colour order kind {kind} n_or {n_or} sweep {sw} replica {r}")


def test_fused_order_strategy_rules(gpu):
    """A sweep without OR passes is the same in either pass order, and
    switching the order re-times the autotuner's pair."""
    out = []
    for fused in (True, False):
        ctx = gpu.Context.pyrochlore(2, n_replicas=4, precision=32, seed=3)
        ctx.define_J1J2J3(1.0, -0.05, 0.05)
        ctx.setup()
        ctx.randomize(1)
        ctx.set_temperatures(np.linspace(0.3, 2.0, 4))
        ctx.set_strategy(3)
        ctx.set_fused_passes(fused)
        acc = ctx.sweep(gpu.LIFTED, 0, 3)
        out.append((acc, [ctx.get_spins(r) for r in range(4)], ctx.sweep_order(0)))
        ctx.close()
    assert np.array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[0][2][0], out[1][2][0]) and np.array_equal(out[0][2][1], out[1][2][1])
    ctx = gpu.Context.pyrochlore(2, n_replicas=4, precision=32, seed=3)
    ctx.define_J1J2J3(1.0, -0.05, 0.05)
    ctx.setup()
    ctx.randomize(1)
    ctx.set_temperatures(np.linspace(0.3, 2.0, 4))
    ctx.sweep(gpu.METROPOLIS, 2, 1)
    assert ctx.strategy() in (3, 4, 6)
    ctx.set_fused_passes(True)
    ctx.sweep(gpu.METROPOLIS, 2, 1)
    assert ctx.strategy() in (3, 4, 6)
    ctx.close()


def test_tile_shape_rules(gpu):
    ctx = gpu.Context.pyrochlore(12, n_replicas=1, precision=32)
    ctx.define_J1J2J3(1.0, 0.1, 0.1)
    ctx.setup()
    assert ctx.tile_shape() == (0, 0, 0)  # 12 is not divisible by 8 on the third axis
    with pytest.raises(gpu.PyroError):
        ctx.set_tile_shape((4, 4, 4))
    with pytest.raises(gpu.PyroError):
        ctx.set_tile_shape((5, 5, 5))
    ctx.close()
    ctx = gpu.Context.pyrochlore(48, n_replicas=1, precision=32)
    ctx.define_J1J2J3(1.0, 0.1, 0.1)
    ctx.setup()
    assert ctx.tile_shape() == (6, 6, 4)
    ctx.set_tile_shape((8, 4, 4))
    assert ctx.tile_shape() == (8, 4, 4)
    site, koff = ctx.sweep_order(0)
    assert np.array_equal(np.sort(site), np.arange(ctx.N))
    ctx.set_tile_shape(None)
    assert ctx.tile_shape() == (6, 6, 4)
    ctx.close()


@pytest.mark.parametrize("L", [1, 2, (1, 2, 3), (2, 1, 4), (3, 2, 1)])
@pytest.mark.parametrize("prec", [64, 32])
def test_small_and_rectangular_lattices(gpu, oracle, L, prec):
    """Lattices of one or two cells along an axis: the bonds wrap onto the
    same cells, so one neighbour can be reached by several displacements
    (each counted, as the reference's shell lists count it).  Fields,
    energy and one sweep of every kind against the oracle."""
    J, Delta, B = (1.0, -0.3, 0.2), 0.6, (0.1, -0.2, 0.05)
    ctx, (o,) = make(gpu, oracle, L, J, Delta, prec, B=B, seed=77)
    S = random_spins(ctx.N, np.random.default_rng(11))
    ctx.set_spins(S)
    o.set_spins(S)
    scale = field_scale(J)
    assert np.max(np.abs(ctx.local_fields() - o.local_fields())) <= TOL[prec] * scale
    assert abs(ctx.energy_per_cell()[0] - o.energy_per_cell()) <= TOL[prec] * 16 * 0.5 * field_scale(J, B)
    if prec == 64:
        ctx.set_strategy(3)
        ctx.set_temperatures([0.7])
        colours = ctx.colouring()
        n_col = ctx.info()["n_colours"]
        for kind, n_or in [(0, 1), (1, 1), (2, 0)]:
            p0 = ctx.pass_count()
            acc = ctx.sweep(kind, n_or, 1)
            want = _oracle_sweeps(o, 0, kind, n_or, 1, 0.7, colours, n_col, 77, p0)
            assert int(acc[0]) == want
            assert np.max(np.abs(ctx.get_spins(0) - o.get_spins())) < 1e-12


def test_invalid_contexts_rejected(gpu):
    with pytest.raises(gpu.PyroError):
        gpu.Context.pyrochlore(0)
    with pytest.raises(gpu.PyroError):
        gpu.Context.pyrochlore(2, n_replicas=0)
    with pytest.raises(gpu.PyroError):
        gpu.Context.pyrochlore(2, precision=16)
    ctx = gpu.Context.pyrochlore(2)
    with pytest.raises(gpu.PyroError):
        ctx.sweep(gpu.METROPOLIS, 1, 1)  # before setup
    ctx.define_J1J2J3(1.0)
    ctx.setup()
    with pytest.raises(gpu.PyroError):
        ctx.sweep(7, 1, 1)
    with pytest.raises(gpu.PyroError):
        ctx.set_temperatures([-1.0])
    ctx.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_tempering_ssf_every_sweep(gpu, prec):
    """Rounds of two sweeps with an SSF sample after each sweep (sample_ssf
    = 2) visit the same states as one-sweep rounds with an exchange after
    every second round: the same swaps and permutation, bit-identical
    contracted SSF sums, and sublattice-resolved sums equal to FP32-MFMA
    rounding although a 3-row buffer is flushed at different points (before
    a round whose two samples would not fit)."""
    out = []
    for per_sweep in (True, False):
        ctx = gpu.Context.pyrochlore(2, n_replicas=6, precision=prec, seed=17)
        ctx.define_J1J2J3(1.0, -0.05, 0.05)
        ctx.setup()
        ctx.randomize(4)
        ctx.pt_init(np.geomspace(2.0, 0.2, 6), 8)
        ctx.ssf_configure(("xx", "zy"), n_slots=6, munu=True, munu_buffer=3)
        if per_sweep:
            ctx.pt_rounds(gpu.LIFTED, 1, 2, 2, 5)
        else:
            ctx.pt_exchange_interval(2)
            ctx.pt_rounds(gpu.LIFTED, 1, 1, True, 10)
        st = ctx.pt_stats()
        ssf = [ctx.ssf_read(s) for s in range(6)]
        munu = [ctx.ssf_read_munu(s, c) for s in range(6) for c in range(2)]
        out.append((st, ctx.pt_permutation(), ssf, munu, [ctx.get_spins(r) for r in range(6)]))
        ctx.close()
    (sa, pa, fa, ma, Sa), (sb, pb, fb, mb, Sb) = out
    assert np.array_equal(pa, pb)
    assert np.array_equal(sa["attempts"], sb["attempts"]) and np.array_equal(sa["accepts"], sb["accepts"])
    assert np.all(sa["n"] == 5) and np.all(sb["n"] == 10)
    for a, b in zip(Sa, Sb):
        assert np.array_equal(a, b)
    for (S1, Q1, n1), (S2, Q2, n2) in zip(fa, fb):
        assert n1 == n2 == 10
        assert np.array_equal(S1, S2) and np.array_equal(Q1, Q2)
    for a, b in zip(ma, mb):
        assert np.abs(a - b).max() <= 2e-5 * 8 * 10


def test_xcd_wait_timeout_is_reported(gpu):
    """The XCD-resident sweep's stage waits are bounded: with the bound set
    to one tick of the device clock a wait times out, the kernel still ends,
    and the next read reports the error instead of returning invalid spins;
    with the bound restored the context sweeps and reads normally again."""
    ctx = gpu.Context.pyrochlore(6, n_replicas=16, precision=32, seed=5)
    ctx.define_J1J2J3(1.0, -0.05, 0.05)
    ctx.setup()
    ctx.randomize(3)
    ctx.set_temperatures(np.geomspace(2.0, 0.1, 16))
    ctx.set_strategy(6)
    ctx.set_wait_timeout(1e-9)
    ctx.sweep(gpu.METROPOLIS, 2, 20, count=False)
    with pytest.raises(gpu.PyroError, match="timed out"):
        ctx.get_spins(0)
    ctx.set_wait_timeout(0.25)
    ctx.randomize(3)
    ctx.sweep(gpu.METROPOLIS, 2, 2, count=False)
    S = ctx.get_spins(0)
    assert np.allclose(np.linalg.norm(S, axis=1), 1.0, atol=1e-5)
    ctx.close()
