"""Host-side checks of libpyro_cmc.so that need no GPU: exported symbols,
the neighbour tables, per-slot couplings, graph colouring and error paths.
The topology is built on the host by pyro_setup_lattice before any device
work, so it can be queried here even when setup then reports no device."""
import re
from pathlib import Path

import numpy as np
import pytest

from helpers import pyro_bond_lists, random_spins

ROOT = Path(__file__).resolve().parent.parent


def _setup_host_only(ctx, pyro):
    try:
        ctx.setup()
    except pyro.PyroError as e:  # no device here: topology is still built
        assert e.code == 4, e


def test_header_symbols_exported(pyro):
    hdr = (ROOT / "include" / "pyro_cmc.h").read_text()
    declared = set(re.findall(r"^(?:int|const char\*)\s+(pyro_[a-z_0-9]+)\s*\(", hdr, re.M))
    assert len(declared) >= 40
    L = pyro.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert declared <= set(pyro.exported_symbols())


def test_create_without_device_is_host_only(pyro):
    ctx = pyro.Context.pyrochlore(2, n_replicas=3, precision=pyro.FP32)
    info = ctx.info()
    assert info["N"] == 128 and info["n_cells"] == 8 and info["R"] == 3
    assert info["n_slot"] == -1  # not set up yet


@pytest.mark.parametrize("J,n_slot,n_col", [((1, 0, 0), 6, 4), ((1, 0.1, 0), 18, 4), ((1, 0.1, 0.2), 30, 16)])
def test_tables_and_colouring(pyro, oracle, J, n_slot, n_col):
    L = 3
    ctx = pyro.Context.pyrochlore(L)
    ctx.define_J1J2J3(*J)
    _setup_host_only(ctx, pyro)
    info = ctx.info()
    assert info["n_slot"] == n_slot and info["n_colours"] == n_col and info["isotropic"] == 1
    o = oracle.Oracle.pyrochlore(L, *J)
    assert np.array_equal(ctx.positions(), o.positions())
    nbr, Jm = ctx.neighbours()
    # field from the product's tables equals the oracle's local_field
    S = random_spins(ctx.N, np.random.default_rng(0))
    h = np.einsum("iKab,iKb->ia", Jm, S[nbr])
    o.set_spins(S)
    assert np.allclose(h, o.local_fields(), rtol=0, atol=1e-13)
    # bonded pairs of non-zero couplings == oracle bond shells with J != 0
    pairs = {(i, int(j)) for i in range(ctx.N) for k, j in enumerate(nbr[i]) if np.any(Jm[i, k])}
    o2 = oracle.Oracle(*oracle.pyro_cell(), L)
    for name, rel, Jc in oracle.j1j2j3_couplings(*J):
        if np.any(Jc):
            o2.define_coupling(name, rel, Jc)
    o2.setup()
    assert pairs == {tuple(p) for p in o2.bonds().tolist()}
    # colouring: no bond inside a colour, and colours partition sublattices
    col = ctx.colouring()
    b = o2.bonds()
    assert np.all(col[b[:, 0]] != col[b[:, 1]])
    assert len(set(col.tolist())) == n_col
    # canonical colours used by the stencil kernels: pyro_sl (4) or sublattice (16)
    sl = np.arange(ctx.N) // L ** 3
    assert np.array_equal(col, sl % 4 if n_col == 4 else sl)


def test_xxz_coupling_matrices(pyro, oracle):
    L = 2
    Delta = 0.35
    ctx = pyro.Context.pyrochlore(L)
    ctx.define_J1J2J3(1.3, -0.2, 0.1, Delta)
    _setup_host_only(ctx, pyro)
    assert ctx.info()["isotropic"] == 0
    nbr, Jm = ctx.neighbours()
    S = random_spins(ctx.N, np.random.default_rng(5))
    h = np.einsum("iKab,iKb->ia", Jm, S[nbr])
    o = oracle.Oracle.pyrochlore(L, 1.3, -0.2, 0.1, Delta)
    o.set_spins(S)
    assert np.allclose(h, o.local_fields(), rtol=0, atol=1e-13)
    # energy from the tables equals the distance-search pair sum
    from helpers import brute_energy_per_cell
    e = 0.5 * (S * h).sum() / L ** 3
    assert e == pytest.approx(brute_energy_per_cell(S, L, 1.3, -0.2, 0.1, Delta), rel=1e-12, abs=1e-12)


def test_square_lattice_needs_residue_colouring(pyro):
    # one site per cell bonded to itself: colours must split by cell parity
    ctx = pyro.Context(np.zeros((1, 3)), [0], [1, 1, 1], [8, 8, 1])
    ctx.define_coupling("J1", [[(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0)]], -np.eye(3))
    ctx.define_coupling("J2", [[(1, 1, 0), (1, -1, 0), (-1, 1, 0), (-1, -1, 0)]], 0.3 * np.eye(3))
    _setup_host_only(ctx, pyro)
    nbr, Jm = ctx.neighbours()
    col = ctx.colouring()
    assert ctx.info()["n_colours"] == 4
    for i in range(ctx.N):
        assert np.all(col[nbr[i]] != col[i])


def test_duplicate_coupling_name_is_logic_error(pyro):
    ctx = pyro.Context.pyrochlore(2)
    ctx.define_coupling("J2", [[(2, 2, 4)]] * 4, np.eye(3))
    with pytest.raises(pyro.PyroError) as e:
        ctx.define_coupling("J2", [[(2, 2, 4)]] * 4, np.eye(3))
    assert e.value.code == 3 and "unique" in str(e.value)


def test_bad_displacement_is_invalid(pyro):
    ctx = pyro.Context.pyrochlore(2)
    ctx.define_coupling("bad", [[(1, 0, 0)]] * 4, np.eye(3))
    with pytest.raises(pyro.PyroError) as e:
        ctx.setup()
    assert e.value.code == 1


def test_missing_kind_list_is_invalid(pyro):
    ctx = pyro.Context.pyrochlore(2)
    ctx.define_coupling("short", [[(0, 2, 2)]], np.eye(3))
    with pytest.raises(pyro.PyroError) as e:
        ctx.setup()
    assert e.value.code == 1


def test_invalid_arguments(pyro):
    with pytest.raises(pyro.PyroError):
        pyro.Context.pyrochlore(2, precision=16)
    with pytest.raises(pyro.PyroError):
        pyro.Context.pyrochlore(2, n_replicas=0)
    ctx = pyro.Context.pyrochlore(2)
    with pytest.raises(pyro.PyroError) as e:
        ctx.sweep()
    assert e.value.code == 2  # before setup
