"""Shared test helpers: configurations, closed forms and brute-force sums.

Everything here is computed independently of both the product and the
oracle (numpy, from positions and the mathematics), so it can pin them.
"""
from __future__ import annotations

import numpy as np

FCC = np.array([[0, 0, 0], [0, 4, 4], [4, 0, 4], [4, 4, 0]])
LINK = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]])


def pyro_positions(L):
    """Site positions in storage order sl*L^3 + (n0*L + n1)*L + n2."""
    out = []
    for f in range(4):
        for p in range(4):
            base = FCC[f] + LINK[p]
            for n0 in range(L):
                for n1 in range(L):
                    for n2 in range(L):
                        out.append(base + 8 * np.array([n0, n1, n2]))
    return np.array(out)


def pyro_kinds(L):
    return np.repeat(np.tile(np.arange(4), 4), L ** 3)


def random_spins(n, rng):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def spiral_spins(pos, Q, L, axis=2):
    """init_spiral_state (cli_bits.hpp:211-222) with integer Q: the phase
    q.r_i with q along `axis`; sin on component axis+1, cos on axis+2."""
    q = 2 * np.pi * Q / (8 * L)
    ph = q * pos[:, axis]
    S = np.zeros((len(pos), 3))
    S[:, (axis + 2) % 3] = np.cos(ph)
    S[:, (axis + 1) % 3] = np.sin(ph)
    return S


def byhand_spiral(pos, qvec):
    """spiral_byhand.cpp:100-107: S_i = (0, cos q.r_i, sin q.r_i)."""
    ph = pos @ np.asarray(qvec, float)
    return np.stack([np.zeros_like(ph), np.cos(ph), np.sin(ph)], 1)


def min_image(d, period):
    return (d + period // 2) % period - period // 2


def pyro_bond_lists(L):
    """All bonded pairs (i < j) of the pyrochlore supercell found by a
    distance search: class 1 (|d|^2 = 8), 2 (24), 3 (32, midpoint on a site),
    4 (32, midpoint empty).  Independent of any displacement table."""
    pos = pyro_positions(L)
    period = 8 * L
    key = {tuple(np.mod(p, period)): i for i, p in enumerate(pos)}
    d = min_image(pos[None, :, :] - pos[:, None, :], period)
    d2 = (d ** 2).sum(-1)
    out = {1: [], 2: [], 3: [], 4: []}
    iu, ju = np.triu_indices(len(pos), 1)
    for i, j in zip(iu, ju):
        dd = d2[i, j]
        if dd == 8:
            out[1].append((i, j))
        elif dd == 24:
            out[2].append((i, j))
        elif dd == 32:
            mid = tuple(np.mod(pos[i] + d[i, j] // 2, period))
            out[3 if mid in key else 4].append((i, j))
    return {k: np.array(v) for k, v in out.items()}, pos


def brute_energy_per_cell(S, L, J1, J2, J3, Delta=1.0, B=(0, 0, 0), bonds=None):
    """Pairwise energy sum_bonds S_i^T J S_j - B.sum S, per cubic cell.
    J1 bonds carry J1[S_mu.S_nu + (Delta-1)(S_mu.z_mu)(S_nu.z_nu)]."""
    if bonds is None:
        bonds, _ = pyro_bond_lists(L)
    kinds = pyro_kinds(L)
    z = LINK / np.sqrt(3.0)
    E = 0.0
    b1 = bonds[1]
    dots = (S[b1[:, 0]] * S[b1[:, 1]]).sum(1)
    zi = (S[b1[:, 0]] * z[kinds[b1[:, 0]]]).sum(1)
    zj = (S[b1[:, 1]] * z[kinds[b1[:, 1]]]).sum(1)
    E += J1 * (dots + (Delta - 1) * zi * zj).sum()
    for k, J in ((2, J2), (3, J3), (4, J3)):
        b = bonds[k]
        if len(b):
            E += J * (S[b[:, 0]] * S[b[:, 1]]).sum()
    E -= np.dot(S.sum(0), np.asarray(B, float))
    return E / L ** 3


def langevin(x):
    return 1.0 / np.tanh(x) - 1.0 / x


def langevin_second(x):
    """<cos^2 theta> for density ~ exp(x cos theta) on the sphere."""
    return 1.0 - 2.0 * langevin(x) / x


def batch_mean_err(series, n_batch=50):
    s = np.asarray(series, float)
    bs = len(s) // n_batch
    m = s[: bs * n_batch].reshape(n_batch, bs).mean(1)
    return s.mean(), m.std(ddof=1) / np.sqrt(n_batch)


def kvec(L, K):
    """Centred wavevector q(K) = 2 pi K' / (8 L) for flat k index K."""
    K0, K1, K2 = K // (L * L), (K // L) % L, K % L
    c = np.array([k if 2 * k < L else k - L for k in (K0, K1, K2)])
    return 2 * np.pi * c / (8 * L)


def direct_ssf(S, pos, L, a, b):
    """S^{ab}(K) = conj(sum_i e^{-iq.r_i} S^a_i) (sum_j e^{-iq.r_j} S^b_j)
    over all K of the L^3 grid (the physical structure factor times N)."""
    out = np.zeros(L ** 3, complex)
    for K in range(L ** 3):
        ph = np.exp(-1j * pos @ kvec(L, K))
        ba = (ph * S[:, a]).sum()
        bb = (ph * S[:, b]).sum()
        out[K] = np.conj(ba) * bb
    return out


def brute_coupling_matrix(L, J1, J2, J3, Delta=1.0):
    """Real-space coupling matrix H[3N][3N] of the supercell from the
    distance-search bond lists, with the energy sum_bonds S_i^T J S_j =
    S^T H S / 2: block (i, j) holds the coupling of bond i -> j and block
    (j, i) its transpose.  Needs L >= 2 (no bond reaches its own image)."""
    bonds, _ = pyro_bond_lists(L)
    kinds = pyro_kinds(L)
    z = LINK / np.sqrt(3.0)
    N = len(kinds)
    H = np.zeros((3 * N, 3 * N))
    for cls, J in ((1, J1), (2, J2), (3, J3), (4, J3)):
        for i, j in bonds[cls]:
            Jm = J * np.eye(3)
            if cls == 1:
                Jm = Jm + J * (Delta - 1) * np.outer(z[kinds[i]], z[kinds[j]])
            H[3 * i:3 * i + 3, 3 * j:3 * j + 3] += Jm
            H[3 * j:3 * j + 3, 3 * i:3 * i + 3] += Jm.T
    return H
