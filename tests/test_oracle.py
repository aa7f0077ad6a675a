"""Pins of the CPU oracle (no GPU).

Each check is chosen so a plausible slip in the restatement fails it:
published known-answer vectors (Philox4x32-10, SplitMix64), a table-free
distance search for the neighbour shells, closed-form energies and structure
factors, brute-force pair sums, symmetry invariants and exact single-spin
statistics (Langevin function).
"""
import numpy as np
import pytest

from helpers import (brute_energy_per_cell, batch_mean_err, byhand_spiral, direct_ssf, langevin,
                     langevin_second, pyro_bond_lists, pyro_positions, random_spins,
                     spiral_spins)

# Philox4x32-10 known-answer vectors of the Random123 distribution (kat_vectors).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(oracle, ctr, key, want):
    assert tuple(int(x) for x in oracle.philox(ctr, key)) == want


def test_splitmix_seeding_known_answer(oracle):
    # Xoroshiro128++(seed) takes its state from SplitMix64(seed); SplitMix64
    # from state 0 first returns 0xE220A8397B1DCDAF (Vigna's reference code).
    # The first Xoroshiro output is rotl(s0 + s1, 17) + s0 of that state.
    s0, s1 = 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4
    mask = (1 << 64) - 1
    t = (s0 + s1) & mask
    first = ((((t << 17) | (t >> 47)) & mask) + s0) & mask
    assert int(oracle.xoroshiro(0, 1)[0]) == first


def test_xoroshiro_uniformity(oracle):
    x = oracle.xoroshiro(12345, 200000).astype(np.float64) / 2.0 ** 64
    assert abs(x.mean() - 0.5) < 5 * np.sqrt(1 / 12 / len(x))
    h, _ = np.histogram(x, bins=16, range=(0, 1))
    chi2 = ((h - len(x) / 16) ** 2 / (len(x) / 16)).sum()
    assert chi2 < 45  # 15 dof, p ~ 1e-4


def _table_bonds(o):
    b = o.bonds()
    return {tuple(sorted(p)) for p in b.tolist()}


@pytest.mark.parametrize("which,couplings", [(1, (1, 0, 0)), (2, (0, 1, 0)), (3, (0, 0, 1))])
def test_shell_tables_match_distance_search(oracle, which, couplings):
    L = 2
    bonds, _ = pyro_bond_lists(L)
    o = oracle.Oracle.pyrochlore(L, *couplings)
    got = _table_bonds(o)
    if which == 3:
        want = {tuple(p) for p in bonds[3].tolist()} | {tuple(p) for p in bonds[4].tolist()}
    else:
        want = {tuple(p) for p in bonds[which].tolist()}
    # zero couplings still define shells in the reference: keep only the
    # requested coupling's bonds by rebuilding with only that coupling
    o2 = oracle.Oracle(*oracle.pyro_cell(), L)
    for name, rel, J in oracle.j1j2j3_couplings(*couplings):
        if np.any(J):
            o2.define_coupling(name, rel, J)
    o2.setup()
    got = _table_bonds(o2)
    assert got == want
    # every bond is listed from both ends
    b = o2.bonds()
    assert len(b) == 2 * len(want)


def test_nn3_split_by_midpoint(oracle):
    bonds, _ = pyro_bond_lists(2)
    for name, cls_ in (("J3a", 3), ("J3b", 4)):
        o = oracle.Oracle(*oracle.pyro_cell(), 2)
        rel = dict((n, r) for n, r, _ in oracle.j1j2j3_couplings(0, 0, 1))[name]
        o.define_coupling(name, rel, np.eye(3))
        o.setup()
        assert _table_bonds(o) == {tuple(p) for p in bonds[cls_].tolist()}


def test_positions_storage_order(oracle):
    o = oracle.Oracle.pyrochlore(3)
    assert np.array_equal(o.positions(), pyro_positions(3))


@pytest.mark.parametrize("J", [(1.0, 0.0, 0.0), (1.0, 0.1, 0.05), (-1.0, -0.2, 0.3)])
def test_ferromagnet_closed_form(oracle, J):
    J1, J2, J3 = J
    B = np.array([0.1, -0.3, 0.2])
    o = oracle.Oracle.pyrochlore(2, J1, J2, J3)
    o.set_field(B)
    m = np.array([1.0, 2.0, -2.0]) / 3.0
    o.set_spins(np.tile(m, (o.N, 1)))
    want = 48 * J1 + 96 * J2 + 96 * J3 - 16 * B @ m
    assert o.energy_per_cell() == pytest.approx(want, rel=1e-13, abs=1e-12)
    h = o.local_fields()
    assert np.allclose(h, (6 * J1 + 12 * J2 + 12 * J3) * m, rtol=0, atol=1e-13)


@pytest.mark.parametrize("Q,axis", [(1, 2), (2, 0), (3, 1)])
def test_spiral_closed_form(oracle, Q, axis):
    L = 4
    J1, J2, J3 = 1.0, 0.15, -0.1
    o = oracle.Oracle.pyrochlore(L, J1, J2, J3)
    S = spiral_spins(o.positions(), Q, L, axis)
    o.set_spins(S)
    q = np.zeros(3)
    q[axis] = 2 * np.pi * Q / (8 * L)
    want = 0.0
    shells = {1: (oracle.NN1, J1), 2: (oracle.NN2, J2), 3: (oracle.NN3A, J3), 4: (oracle.NN3B, J3)}
    for table, Jd in shells.values():
        for p in range(4):
            for d in table[p]:
                want += 2 * Jd * np.cos(q @ np.array(d))
    assert o.energy_per_cell() == pytest.approx(want, rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("Delta,B", [(1.0, (0, 0, 0)), (0.4, (0.2, 0.0, -0.5)), (-1.3, (0, 1, 0))])
def test_energy_matches_brute_force_pairs(oracle, Delta, B):
    L = 2
    rng = np.random.default_rng(7)
    S = random_spins(16 * L ** 3, rng)
    J1, J2, J3 = 1.0, -0.2, 0.15
    o = oracle.Oracle.pyrochlore(L, J1, J2, J3, Delta)
    o.set_field(np.array(B, float))
    o.set_spins(S)
    want = brute_energy_per_cell(S, L, J1, J2, J3, Delta, B)
    assert o.energy_per_cell() == pytest.approx(want, rel=1e-12, abs=1e-12)


def test_energy_rotation_invariant_and_or_preserves_energy(oracle):
    L = 2
    rng = np.random.default_rng(3)
    o = oracle.Oracle.pyrochlore(L, 1.0, 0.1, -0.05)
    S = random_spins(o.N, rng)
    o.set_spins(S)
    e0 = o.energy_per_cell()
    A = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    o.set_spins(S @ A.T)
    assert o.energy_per_cell() == pytest.approx(e0, rel=1e-12, abs=1e-12)
    # one overrelaxation pass (counter mode) keeps the energy
    colours = np.arange(o.N) // (o.N // 16)  # one colour per sublattice
    o.set_spins(S)
    o.pass_ctr(3, 1.0, colours, 16, seed=5, replica=0, pass_index=0)
    assert not np.allclose(o.get_spins(), S)
    assert o.energy_per_cell() == pytest.approx(e0, rel=1e-12, abs=1e-12)


def _independent_spins(oracle, B):
    o = oracle.Oracle(*oracle.pyro_cell(), 2, seed=11)
    o.setup()  # no couplings: independent spins in the field
    o.set_field(np.asarray(B, float))
    return o


@pytest.mark.parametrize("kind,n_or,T", [(0, 0, 0.5), (0, 1, 0.5), (1, 0, 0.7), (1, 1, 0.4)])
def test_reference_sampler_langevin(oracle, kind, n_or, T):
    """With J = 0 every spin is independent: <S.B^> = L(|B|/T) exactly.
    (The lifted-rotation move only steps the polar angle by a fixed delta
    about a fixed field, so it is not ergodic here; it is checked on a
    coupled model below.)"""
    B = np.array([0.3, -0.6, 0.6])
    Bn = np.linalg.norm(B)
    o = _independent_spins(oracle, B)
    o.set_spins(np.tile([1.0, 0, 0], (o.N, 1)))
    for _ in range(200):
        o.sweep_ref(kind, T, n_or)
    m1, m2 = [], []
    for _ in range(6000):
        o.sweep_ref(kind, T, n_or)
        c = o.get_spins() @ (B / Bn)
        m1.append(c.mean())
        m2.append((c ** 2).mean())
    x = Bn / T
    for series, want in ((m1, langevin(x)), (m2, langevin_second(x))):
        mean, err = batch_mean_err(series)
        assert abs(mean - want) < 4.5 * err + 1e-4, (mean, want, err)


def test_ssf_direct_matches_physical_definition(oracle):
    L = 2
    rng = np.random.default_rng(1)
    o = oracle.Oracle.pyrochlore(L)
    S = random_spins(o.N, rng)
    o.set_spins(S)
    corr = ("xx", "yz", "zy")
    got = o.ssf_sample(corr)
    pos = o.positions()
    for c, (a, b) in enumerate([(0, 0), (1, 2), (2, 1)]):
        want = direct_ssf(S, pos, L, a, b)
        assert np.allclose(got[c, :, 0] + 1j * got[c, :, 1], want, rtol=0, atol=1e-10 * o.N)


def test_ssf_ferromagnet_and_spiral_closed_forms(oracle):
    L = 4
    o = oracle.Oracle.pyrochlore(L)
    N = o.N
    m = np.array([2.0, -1.0, 2.0]) / 3.0
    o.set_spins(np.tile(m, (N, 1)))
    out = o.ssf_sample(("xx", "xz", "yz"))
    for c, (a, b) in enumerate([(0, 0), (0, 2), (1, 2)]):
        v = out[c, :, 0] + 1j * out[c, :, 1]
        assert v[0] == pytest.approx(N * N * m[a] * m[b], rel=1e-12)
        assert np.abs(v[1:]).max() < 1e-9 * N * N
    # spiral_byhand state along z with Q = 1: weight N^2/4 at K = (0,0,+-1)
    S = byhand_spiral(o.positions(), [0, 0, 2 * np.pi / (8 * L)])
    o.set_spins(S)
    out = o.ssf_sample(("xx", "yy", "zz", "yz"))
    Kp, Km = 1, L - 1  # flat index (0,0,1) and (0,0,L-1)
    yy = out[1, :, 0] + 1j * out[1, :, 1]
    zz = out[2, :, 0] + 1j * out[2, :, 1]
    yz = out[3, :, 0] + 1j * out[3, :, 1]
    assert np.abs(out[0]).max() < 1e-9 * N * N
    for K in (Kp, Km):
        assert yy[K] == pytest.approx(N * N / 4, rel=1e-12)
        assert zz[K] == pytest.approx(N * N / 4, rel=1e-12)
    assert yz[Kp] == pytest.approx(-1j * N * N / 4, rel=1e-12)
    assert yz[Km] == pytest.approx(+1j * N * N / 4, rel=1e-12)
    mask = np.ones(L ** 3, bool)
    mask[[Kp, Km]] = False
    assert np.abs(yy[mask]).max() < 1e-9 * N * N


def test_ssf_hermitian_and_parseval(oracle):
    L = 3
    rng = np.random.default_rng(4)
    o = oracle.Oracle.pyrochlore(L)
    S = random_spins(o.N, rng)
    o.set_spins(S)
    out, munu = o.ssf_sample(("xy", "yx"), munu=True)
    xy = out[0, :, 0] + 1j * out[0, :, 1]
    yx = out[1, :, 0] + 1j * out[1, :, 1]
    assert np.allclose(yx, np.conj(xy), atol=1e-10 * o.N)
    ft = o.spin_ft()
    A = ft[..., 0] + 1j * ft[..., 1]
    # Parseval per (component, sublattice): sum_K |A|^2 = L^3 sum_n |S|^2
    for a in range(3):
        for mu in range(16):
            lhs = (np.abs(A[a, mu]) ** 2).sum()
            rhs = L ** 3 * (S[mu * L ** 3:(mu + 1) * L ** 3, a] ** 2).sum()
            assert lhs == pytest.approx(rhs, rel=1e-12)
    # mu-nu resolved layout: C_{mu nu} = conj(A^x_mu) A^y_nu
    C = munu[0, ..., 0] + 1j * munu[0, ..., 1]
    want = np.conj(A[0].T)[:, :, None] * A[1].T[:, None, :]
    assert np.allclose(C, want, atol=1e-10 * o.N)


def test_exchange_rule(oracle):
    # temper.cpp:456-462: arg = (1/T_s - 1/T_{s+1})(E_s - E_{s+1}) >= 0 always swaps
    T = np.array([2.0, 1.0, 0.5, 0.25])
    e = np.array([-1.0, -2.0, -3.0, -4.0])  # already ordered: hotter has higher E
    att = np.zeros(4, np.uint64)
    acc = np.zeros(4, np.uint64)
    rep = oracle.exchange(T, e, 8.0, np.arange(4), 0, 1, att, acc)
    # arg = (0.5 - 1)(-8 + 16) = -4 < 0 for pair (0,1) -> swap with prob e^-4
    # flipped energies make every arg positive: always swap
    att[:] = 0
    acc[:] = 0
    rep = oracle.exchange(T, e[::-1].copy(), 8.0, np.arange(4), 0, 1, att, acc)
    assert list(rep) == [1, 0, 3, 2] and list(acc[[0, 2]]) == [1, 1]
    # odd rounds pair (1,2)
    att[:] = 0
    acc[:] = 0
    rep = oracle.exchange(T, e[::-1].copy(), 8.0, np.arange(4), 1, 1, att, acc)
    assert list(rep) == [0, 2, 1, 3] and int(att[1]) == 1


# Square-lattice J1-J2 model of the reference's registered test
# (bench_lifted.cpp:37-52): one site per cell, L x L x 1 supercell.
SQUARE_J1 = [[(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0)]]
SQUARE_J2 = [[(1, 1, 0), (1, -1, 0), (-1, 1, 0), (-1, -1, 0)]]


def square_oracle(oracle, L, J1, J2, seed):
    o = oracle.Oracle(np.zeros((1, 3), np.int32), np.zeros(1, np.int32), np.ones(3, np.int32), (L, L, 1), seed)
    o.define_coupling("J1", SQUARE_J1, J1 * np.eye(3))
    o.define_coupling("J2", SQUARE_J2, J2 * np.eye(3))
    o.setup()
    return o


@pytest.mark.parametrize("kind", [1, 2])
def test_bench_lifted_consistency(oracle, kind):
    """bench_lifted.cpp:138-273 defaults: J1=-1, J2=0.3, T=0.5, 8x8; lifted
    and standard Metropolis agree on <E>, <E^2> per spin within
    0.1 min(|J1|,|J2|) (the reference's pass criterion, :200, :260-265)."""
    J1, J2, T, L = -1.0, 0.3, 0.5, 8
    tol = 0.1 * min(abs(J1), abs(J2))
    rng = np.random.default_rng(2718281828)
    res = []
    for k, seed in ((0, 11), (kind, 12)):
        o = square_oracle(oracle, L, J1, J2, seed)
        o.set_spins(random_spins(o.N, rng))
        for _ in range(2000):
            o.sweep_ref(k, T, 0)
        es = []
        for _ in range(20000):
            o.sweep_ref(k, T, 0)
            es.append(o.energy_per_cell())
        es = np.array(es)
        res.append((es.mean(), (es ** 2).mean()))
    assert abs(res[0][0] - res[1][0]) < tol
    assert abs(res[0][1] - res[1][1]) < tol
