"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never loads it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
_vp = C.c_void_p

_SIGS = {
    "orc_create": (_vp, [C.c_int, _i32p, _i32p, _i32p, _i32p, C.c_uint64]),
    "orc_destroy": (None, [_vp]),
    "orc_num_sites": (C.c_int64, [_vp]),
    "orc_define_coupling": (C.c_int, [_vp, C.c_char_p, C.c_int, _i32p, _i32p, _f64p]),
    "orc_setup": (C.c_int, [_vp]),
    "orc_set_coupling": (C.c_int, [_vp, C.c_char_p, _f64p]),
    "orc_set_field": (None, [_vp, _f64p]),
    "orc_set_T_ref": (None, [_vp, C.c_double]),
    "orc_set_spins": (None, [_vp, _f64p]),
    "orc_get_spins": (None, [_vp, _f64p]),
    "orc_set_dir": (None, [_vp, _i8p]),
    "orc_get_dir": (None, [_vp, _i8p]),
    "orc_positions": (None, [_vp, _i32p]),
    "orc_bonds": (C.c_int64, [_vp, _i64p, C.c_int64]),
    "orc_local_fields": (None, [_vp, _f64p]),
    "orc_energy_per_cell": (C.c_double, [_vp]),
    "orc_sweep_ref": (C.c_uint64, [_vp, C.c_int, C.c_double, C.c_int]),
    "orc_pass_ctr": (C.c_uint64, [_vp, C.c_int, C.c_double, _i32p, C.c_int, C.c_uint64, C.c_uint32, C.c_uint64]),
    "orc_randomize_ctr": (None, [_vp, C.c_uint64, C.c_uint32]),
    "orc_sweep_order": (C.c_uint64, [_vp, C.c_int, C.c_double, _i32p, _i32p, C.c_int64, C.c_uint64, C.c_uint32, C.c_uint64]),
    "orc_pass_ctr_m": (C.c_uint64, [_vp, C.c_int, C.c_double, _i32p, C.c_int, C.c_uint64, C.c_uint32, C.c_uint64,
                                    _vp]),
    "orc_sweep_order_m": (C.c_uint64, [_vp, C.c_int, C.c_double, _i32p, _i32p, C.c_int64, C.c_uint64, C.c_uint32,
                                       C.c_uint64, _vp, C.c_int]),
    "orc_philox": (None, [_u32p, _u32p, _u32p]),
    "orc_xoroshiro": (None, [C.c_uint64, C.c_int, _u64p]),
    "orc_exchange": (None, [C.c_int, _f64p, _f64p, C.c_double, _i32p, C.c_uint64, C.c_uint64, _u64p, _u64p]),
    "orc_spin_ft": (None, [_vp, _f64p]),
    "orc_ssf_sample": (None, [_vp, C.c_int, _i32p, _i32p, _f64p, _vp]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        L = C.CDLL(str(LIB))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


# ---------------------------------------------------------------------------
# Pyrochlore geometry restated from pyrochlore_geometry.hpp (tables at
# :16-30, :110-149, :155-190, :199-208; couplings cli_bits.hpp:168-200).
# ---------------------------------------------------------------------------
R_FCC = [(0, 0, 0), (0, 4, 4), (4, 0, 4), (4, 4, 0)]
LINK = [(1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)]

NN1 = [
    [(0, -2, -2), (-2, 0, -2), (-2, -2, 0), (0, 2, 2), (2, 0, 2), (2, 2, 0)],
    [(-2, 2, 0), (-2, 0, 2), (0, 2, 2), (2, -2, 0), (2, 0, -2), (0, -2, -2)],
    [(0, -2, 2), (2, 0, 2), (2, -2, 0), (0, 2, -2), (-2, 0, -2), (-2, 2, 0)],
    [(2, 2, 0), (2, 0, -2), (0, 2, -2), (-2, -2, 0), (-2, 0, 2), (0, -2, 2)],
]
NN2 = [
    [(-2, 4, 2), (-2, 2, 4), (4, -2, 2), (2, -2, 4), (4, 2, -2), (2, 4, -2),
     (2, -4, -2), (2, -2, -4), (-4, 2, -2), (-2, 2, -4), (-4, -2, 2), (-2, -4, 2)],
    [(2, -4, 2), (4, -2, 2), (2, 2, -4), (4, 2, -2), (-2, -2, -4), (-2, -4, -2),
     (-2, 4, -2), (-4, 2, -2), (-2, -2, 4), (-4, -2, 2), (2, 2, 4), (2, 4, 2)],
    [(2, 4, -2), (2, 2, -4), (-4, -2, -2), (-2, -2, -4), (-4, 2, 2), (-2, 4, 2),
     (-2, -4, 2), (-2, -2, 4), (4, 2, 2), (2, 2, 4), (4, -2, -2), (2, -4, -2)],
    [(-2, -4, -2), (-4, -2, -2), (-2, 2, 4), (-4, 2, 2), (2, -2, 4), (2, -4, 2),
     (2, 4, 2), (4, 2, 2), (2, -2, -4), (4, -2, -2), (-2, 2, -4), (-2, 4, -2)],
]
NN3A = [
    [(0, 4, 4), (4, 0, 4), (4, 4, 0), (0, -4, -4), (-4, 0, -4), (-4, -4, 0)],
    [(4, -4, 0), (4, 0, -4), (0, -4, -4), (-4, 4, 0), (-4, 0, 4), (0, 4, 4)],
    [(0, 4, -4), (-4, 0, -4), (-4, 4, 0), (0, -4, 4), (4, 0, 4), (4, -4, 0)],
    [(-4, -4, 0), (-4, 0, 4), (0, -4, 4), (4, 4, 0), (4, 0, -4), (0, 4, -4)],
]
NN3B = [
    [(-4, 4, 0), (4, 0, -4), (0, -4, 4), (4, -4, 0), (-4, 0, 4), (0, 4, -4)],
    [(0, -4, 4), (-4, 0, -4), (4, 4, 0), (0, 4, -4), (4, 0, 4), (-4, -4, 0)],
    [(4, 4, 0), (-4, 0, 4), (0, -4, -4), (-4, -4, 0), (4, 0, -4), (0, 4, 4)],
    [(0, -4, -4), (4, 0, 4), (-4, 4, 0), (0, 4, 4), (-4, 0, -4), (4, -4, 0)],
]
NN1_PAIR = {
    (0, 1): [[(0, -2, -2), (0, 2, 2)], [(0, 2, 2), (0, -2, -2)], [], []],
    (0, 2): [[(-2, 0, -2), (2, 0, 2)], [], [(2, 0, 2), (-2, 0, -2)], []],
    (0, 3): [[(-2, -2, 0), (2, 2, 0)], [], [], [(2, 2, 0), (-2, -2, 0)]],
    (1, 2): [[], [(-2, 2, 0), (2, -2, 0)], [(2, -2, 0), (-2, 2, 0)], []],
    (1, 3): [[], [(-2, 0, 2), (2, 0, -2)], [], [(2, 0, -2), (-2, 0, 2)]],
    (2, 3): [[], [], [(0, -2, 2), (0, 2, -2)], [(0, 2, -2), (0, -2, 2)]],
}


def local_z(p):
    s3 = 1.732050807568877293527446341505872366942805253810380628055
    return np.array([(1.0 / s3) * c for c in LINK[p]])


def pyro_cell():
    pos = np.array([[R_FCC[f][a] + LINK[p][a] for a in range(3)] for f in range(4) for p in range(4)], np.int32)
    kind = np.array([p for f in range(4) for p in range(4)], np.int32)
    return pos, kind, np.array([8, 8, 8], np.int32)


def j1j2j3_couplings(J1, J2, J3, Delta=1.0):
    """The nine couplings of build_J1J2J3_h (cli_bits.hpp:168-200)."""
    out = []
    for (mu, nu), rel in NN1_PAIR.items():
        zm, zn = local_z(mu), local_z(nu)
        ising = np.array([[zn[b] * zm[a] for b in range(3)] for a in range(3)])
        J = J1 * (Delta - 1) * ising + J1 * np.eye(3)
        out.append((f"J1_{mu}{nu}", rel, J))
    out.append(("J2", NN2, J2 * np.eye(3)))
    out.append(("J3a", NN3A, J3 * np.eye(3)))
    out.append(("J3b", NN3B, J3 * np.eye(3)))
    return out


# ---------------------------------------------------------------------------
# Luttinger-Tisza matrix: bonds as in luttinger_tisza.cpp:31-81, the k sweep
# and first-strict-minimum search as in :147-181.  The q representative is
# the centred index K' (K < L/2 ? K : K - L); eigenvalues do not depend on
# it, eigenvectors change by a sublattice-diagonal phase.
# ---------------------------------------------------------------------------
def lt_bonds(J1, J2=0.0, J3=0.0, Delta=1.0):
    """Directed bonds (alpha, beta, d, J) of the 16-sublattice cubic cell:
    for each sublattice alpha and displacement v of its kind's shell, beta =
    the sublattice at r_alpha + v modulo the edge 8; nn1 carries
    J1 [I + (Delta - 1) z_alpha z_beta^T], the other shells J I; zero
    couplings are skipped (luttinger_tisza.cpp:46-47)."""
    pos, _, _ = pyro_cell()
    where = {tuple(int(c) % 8 for c in p): s for s, p in enumerate(pos)}
    out = []
    for j, table, xxz in ((J1, NN1, True), (J2, NN2, False), (J3, NN3A, False), (J3, NN3B, False)):
        if j == 0.0:
            continue
        for a in range(16):
            for v in table[a % 4]:
                b = where[tuple(int(pos[a][c] + v[c]) % 8 for c in range(3))]
                J = np.eye(3)
                if xxz and Delta != 1.0:
                    J = J + (Delta - 1.0) * np.outer(local_z(a % 4), local_z(b % 4))
                out.append((a, b, np.array(v, np.float64), j * J))
    return out


def lt_q(K, L):
    """q = 2 pi K' / (8 L) per axis for the integer index K."""
    K = np.asarray(K)
    L = np.asarray(L)
    Kc = np.where(2 * K < L, K, K - L)
    return 2 * np.pi * Kc / (8.0 * L)


def lt_k_grid(L):
    """All k indices of an L0 x L1 x L2 grid in flat order (K0 L1 + K1) L2 + K2."""
    L = [L] * 3 if np.isscalar(L) else list(L)
    K = np.stack(np.meshgrid(*[np.arange(n) for n in L], indexing="ij"), -1).reshape(-1, 3)
    return K, np.array(L)


def lt_matrices(q, bonds):
    """M(k)[48][48] for q[n_k][3]: block (alpha, beta) += J e^{i q.d}."""
    q = np.atleast_2d(q)
    M = np.zeros((len(q), 48, 48), complex)
    for a, b, d, J in bonds:
        M[:, 3 * a:3 * a + 3, 3 * b:3 * b + 3] += J[None] * np.exp(1j * (q @ d))[:, None, None]
    return M


def lt_spectrum(L, J1, J2=0.0, J3=0.0, Delta=1.0):
    """(lambda_min[n_k], k_star[3], eigvec[48]) with k_star the first k of
    the strict minimum (luttinger_tisza.cpp:150-181)."""
    K, Ls = lt_k_grid(L)
    bonds = lt_bonds(J1, J2, J3, Delta)
    lam = np.empty(len(K))
    best, vec, k_star = np.inf, None, None
    for k0 in range(0, len(K), 4096):
        M = lt_matrices(lt_q(K[k0:k0 + 4096], Ls), bonds)
        w, v = np.linalg.eigh(M)
        lam[k0:k0 + len(M)] = w[:, 0]
        for i in range(len(M)):
            if w[i, 0] < best:
                best, vec, k_star = w[i, 0], v[i, :, 0], K[k0 + i]
    return lam, np.asarray(k_star), vec


class Oracle:
    """One replica of the CPU restatement."""

    def __init__(self, sl_pos, sl_kind, cell, L, seed=0):
        self.sl_pos = np.ascontiguousarray(sl_pos, np.int32).reshape(-1, 3)
        self.sl_kind = np.ascontiguousarray(sl_kind, np.int32)
        self.cell = np.ascontiguousarray(cell, np.int32)
        self.L = np.ascontiguousarray([L] * 3 if np.isscalar(L) else L, np.int32)
        self.h = lib().orc_create(len(self.sl_kind), self.sl_pos.ravel(), self.sl_kind, self.cell, self.L, seed)
        self.N = int(lib().orc_num_sites(self.h))
        self.n_sl = len(self.sl_kind)
        self.n_cells = int(np.prod(self.L))

    @classmethod
    def pyrochlore(cls, L, J1=1.0, J2=0.0, J3=0.0, Delta=1.0, seed=0):
        pos, kind, cell = pyro_cell()
        o = cls(pos, kind, cell, L, seed)
        for name, rel, J in j1j2j3_couplings(J1, J2, J3, Delta):
            o.define_coupling(name, rel, J)
        o.setup()
        return o

    def __del__(self):
        try:
            lib().orc_destroy(self.h)
        except Exception:
            pass

    def define_coupling(self, name, rel, J):
        counts = np.array([len(v) for v in rel], np.int32)
        flat = np.array([x for v in rel for d in v for x in d], np.int32) if counts.sum() else np.zeros(3, np.int32)
        rc = lib().orc_define_coupling(self.h, name.encode(), len(rel), counts, flat,
                                       np.ascontiguousarray(J, np.float64).ravel())
        if rc != 0:
            raise ValueError("Coupling names must be unique")

    def set_coupling(self, name, J):
        """New 3x3 matrix for an existing coupling (no new setup)."""
        if lib().orc_set_coupling(self.h, name.encode(), np.ascontiguousarray(J, np.float64).ravel()) != 0:
            raise KeyError(name)

    def set_J1J2J3(self, J1, J2=0.0, J3=0.0, Delta=1.0):
        """The nine build_J1J2J3_h matrices of another model on an oracle
        made by pyrochlore() (the zero couplings are defined there too)."""
        for name, _, J in j1j2j3_couplings(J1, J2, J3, Delta):
            self.set_coupling(name, J)

    def setup(self):
        if lib().orc_setup(self.h) != 0:
            raise ValueError("displacement does not land on a site")
        return self

    def set_field(self, B):
        lib().orc_set_field(self.h, np.ascontiguousarray(B, np.float64))

    def set_T_ref(self, T_ref):
        lib().orc_set_T_ref(self.h, float(T_ref))

    def set_spins(self, S):
        lib().orc_set_spins(self.h, np.ascontiguousarray(S, np.float64).ravel())

    def get_spins(self):
        S = np.zeros((self.N, 3))
        lib().orc_get_spins(self.h, S.ravel())
        return S

    def set_dir(self, d):
        lib().orc_set_dir(self.h, np.ascontiguousarray(d, np.int8))

    def get_dir(self):
        d = np.zeros(self.N, np.int8)
        lib().orc_get_dir(self.h, d)
        return d

    def positions(self):
        p = np.zeros((self.N, 3), np.int32)
        lib().orc_positions(self.h, p.ravel())
        return p

    def bonds(self):
        n = lib().orc_bonds(self.h, np.zeros(2, np.int64), 0)
        a = np.zeros(2 * n, np.int64)
        lib().orc_bonds(self.h, a, n)
        return a.reshape(n, 2)

    def local_fields(self):
        h = np.zeros((self.N, 3))
        lib().orc_local_fields(self.h, h.ravel())
        return h

    def energy_per_cell(self):
        return float(lib().orc_energy_per_cell(self.h))

    def sweep_ref(self, kind, T, n_or=1):
        return int(lib().orc_sweep_ref(self.h, kind, float(T), n_or))

    def pass_ctr(self, kind, T, colour_of_site, n_colours, seed, replica, pass_index, info=None):
        """Counter-mode pass.  info (float64[N][4], optional) receives each
        site's conditioning record: the distance of its closest decision
        test from the threshold (+inf where it decides nothing), |h_loc|,
        the sensitivity |dS'|/|dh| of its new spin to the field (0 where the
        update does not follow h) and the amplification of its own rounding."""
        p = None if info is None else info.ctypes.data_as(C.c_void_p)
        return int(lib().orc_pass_ctr_m(self.h, kind, float(T), np.ascontiguousarray(colour_of_site, np.int32),
                                        int(n_colours), int(seed), int(replica), int(pass_index), p))

    def sweep_order(self, first_kind, T, site, pass_offset, seed, replica, pass0, info=None, shared=False):
        """Counter-mode updates in an explicit order; info (float64[n][4],
        optional) receives the conditioning record of every update (see
        pass_ctr).  shared: a site's OR passes take the next uniforms of the
        sweep's blocks (the product's fused passes) instead of a block each."""
        site = np.ascontiguousarray(site, np.int32)
        koff = np.ascontiguousarray(pass_offset, np.int32)
        p = None if info is None else info.ctypes.data_as(C.c_void_p)
        return int(lib().orc_sweep_order_m(self.h, int(first_kind), float(T), site, koff, len(site),
                                           int(seed), int(replica), int(pass0), p, int(bool(shared))))

    def randomize_ctr(self, seed, replica=0):
        lib().orc_randomize_ctr(self.h, int(seed), int(replica))

    def spin_ft(self):
        a = np.zeros((3, self.n_sl, self.n_cells, 2))
        lib().orc_spin_ft(self.h, a.ravel())
        return a

    def ssf_sample(self, correlators=("xx", "yy", "zz"), munu=False):
        ax = {"x": 0, "y": 1, "z": 2}
        al = np.array([ax[c[0]] for c in correlators], np.int32)
        be = np.array([ax[c[1]] for c in correlators], np.int32)
        out = np.zeros((len(al), self.n_cells, 2))
        m = np.zeros((len(al), self.n_cells, self.n_sl, self.n_sl, 2)) if munu else None
        lib().orc_ssf_sample(self.h, len(al), al, be, out.ravel(),
                             None if m is None else m.ctypes.data_as(C.c_void_p))
        return (out, m) if munu else out


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def xoroshiro(seed, n):
    out = np.zeros(n, np.uint64)
    lib().orc_xoroshiro(int(seed), int(n), out)
    return out


def exchange(T_slot, e, n_cells, rep_of, rnd, swap_seed, att, accp):
    rep = np.ascontiguousarray(rep_of, np.int32).copy()
    lib().orc_exchange(len(T_slot), np.ascontiguousarray(T_slot, np.float64), np.ascontiguousarray(e, np.float64),
                       float(n_cells), rep, int(rnd), int(swap_seed), att, accp)
    return rep
