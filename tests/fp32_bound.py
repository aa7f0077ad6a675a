"""First-order FP32 error bounds for sweep trajectories.

An FP32 sweep and the FP64 oracle's replay of the same visiting order and
uniforms differ by rounding, which each update amplifies according to its
conditioning, and which moves on to later updates through the local field.
From the oracle's per-update conditioning record (pyoracle pass_ctr /
sweep_order `info`: decision margin, |h|, field sensitivity |dS'/dh|,
rounding amplification) this propagates an absolute error bound E_i per
site in visiting order:

    dh   = EPS_FIELD * scale + sqrt(sum_j (J_ij E_j)^2)      (bonded j)
    E_i += |dS'/dh| * dh + EPS_ROUND * amplification + EPS_UPDATE

A decision may flip when the field perturbation that would flip it (the
record's margin) is within 2 (dh + |h| E_i) + EPS_FIELD * scale.  One flip changes a neighbour's field by O(J), which at
these temperatures flips its decisions with O(1) probability, so no
first-order bound survives it: the flipped site and every later update of a
site bonded to one already reached (its downstream cone) are exempt
("tainted").  Updates are processed stage by stage: a stage holds updates
with no bond between them, so each is vectorised.  The test then requires
every untainted site's deviation to be within FP32_SITE_TOL + SAFETY * E_i.
"""
from __future__ import annotations

import numpy as np

FP32_SITE_TOL = 5e-5
EPS_FIELD = 3e-7   # relative FP32 error of a 30-term field sum
EPS_ROUND = 2e-7   # a few FP32 ulps
EPS_UPDATE = 1e-6  # one FP32 update, hardware transcendentals included
SAFETY = 10.0


def _group(koff, fused):
    """Stage group of each update: with fused passes the first kind and the
    first OR pass share group 0 and OR pass k >= 2 is group k - 1; else
    pass k is group k."""
    return np.maximum(koff - 1, 0) if fused else koff


def tiled_stages(ctx, site, koff, fused=False):
    """Stage key of every update of a tiled sweep's order (pyro_sweep_order
    under the tiled strategy): (checkerboard phase, pass, colour), or with
    fused passes (phase, stage group, colour, pass).  Tiles of one phase
    share no bond and a colour class has no bond inside it, so updates of
    one key are independent, and taking the keys in increasing order
    preserves every dependency of the tile-by-tile order."""
    T = np.array(ctx.tile_shape())
    L = np.array(ctx.Ls)
    nc = int(np.prod(L))
    cell = site % nc
    n = np.stack([cell // (L[1] * L[2]), (cell // L[2]) % L[1], cell % L[2]], 1)
    par = (n // T) % 2
    phase = par[:, 0] + 2 * par[:, 1] + 4 * par[:, 2]
    n_col = ctx.info()["n_colours"]
    colour = ctx.colouring()[site]
    n_pass = int(koff.max()) + 1
    grp = _group(koff, fused)
    return ((phase.astype(np.int64) * n_pass + grp) * n_col + colour) * n_pass + koff


def colour_stages(ctx, site, koff, fused=False):
    """Stage key of every update of a per-colour sweep's order
    (pyro_sweep_order under strategy 3): (pass, colour), or with fused
    passes (stage group, colour, pass)."""
    n_col = ctx.info()["n_colours"]
    colour = ctx.colouring()[site]
    n_pass = int(koff.max()) + 1
    return (_group(koff, fused).astype(np.int64) * n_col + colour) * n_pass + koff


def padded_adjacency(o, J=None):
    """Bonded neighbours of every site (oracle bond shells) as an [N][D]
    array padded with N (a dummy site whose bound is 0), and the coupling
    magnitude of every entry.  An oracle made by Oracle.pyrochlore lists a
    site's bonds by coupling in definition order: the 6 nn1 bonds (J1
    [I + (Delta - 1) z z^T], |J| up to |J1| max(1, |Delta|)), then 12 nn2
    (J2) and 12 nn3 (J3); J = (J1, J2, J3, Delta) gives those weights, else
    every entry weighs 1."""
    pairs = o.bonds()  # sorted by the first index (the oracle lists them site by site)
    counts = np.bincount(pairs[:, 0], minlength=o.N)
    D = int(counts.max())
    start = np.concatenate([[0], np.cumsum(counts)])
    nb = np.full((o.N, D), o.N, np.int64)
    col = np.arange(len(pairs)) - start[pairs[:, 0]]
    nb[pairs[:, 0], col] = pairs[:, 1]
    w = np.ones((o.N, D))
    if J is not None:
        J1, J2, J3 = abs(J[0]) * max(1.0, abs(J[3]) if len(J) > 3 else 1.0), abs(J[1]), abs(J[2])
        row = np.array([J1] * 6 + [J2] * 12 + [J3] * 12 + [max(J1, J2, J3)] * max(0, D - 30))[:D]
        w = np.broadcast_to(row, (o.N, D)) if np.all(counts == 30) else np.full((o.N, D), max(J1, J2, J3))
    return nb, w


def propagate(order_sites, stage, info, adj, T, scale, N):
    """(E, tainted, seeds): error bound per site and the exempt sites after
    the updates order_sites (with stage keys: updates of one key share no
    bond, keys are processed in increasing order) given their conditioning
    records info[n][4] and adj = padded_adjacency(...); seeds = the number
    of possibly flipped decisions."""
    nb, w = adj
    E = np.zeros(N + 1)
    taint = np.zeros(N + 1, bool)
    order_sites = np.asarray(order_sites, np.int64)
    keys = np.unique(stage)
    by_key = np.argsort(stage, kind="stable")
    bounds = np.searchsorted(stage[by_key], np.concatenate([keys, [np.iinfo(np.int64).max]]))
    seeds = 0
    for a, b in zip(bounds[:-1], bounds[1:]):
        ks = by_key[a:b]
        i = order_sites[ks]
        margin, hn, Ah, Ar = info[ks, 0], info[ks, 1], info[ks, 2], info[ks, 3]
        En = np.where(taint[nb[i]], 0.0, E[nb[i]]) * w[i]
        dh = EPS_FIELD * scale + np.sqrt((En ** 2).sum(1))
        seed = (margin < 2.0 * (dh + hn * E[i]) + EPS_FIELD * scale) & ~taint[i]
        seeds += int(seed.sum())
        taint[i] |= seed | taint[nb[i]].any(1)
        E[i] = np.minimum(2.0, E[i] + Ah * dh + EPS_ROUND * Ar + EPS_UPDATE)
    return E[:N], taint[:N], seeds


def check(gpu_S, orc_S, E, taint, seeds, n_updates, label=""):
    """Every untainted site within FP32_SITE_TOL + SAFETY * E, and possibly
    flipped decisions rare (at most 1 in 100 updates, a handful on small
    lattices: candidates by the bound, which is conservative, not observed
    flips); returns the largest ratio of deviation to allowance (printed
    for calibration).  The tainted cones may cover much of a large lattice
    (one flip at these temperatures flips downstream decisions with O(1)
    probability), so their size is printed, not bounded."""
    diff = np.abs(gpu_S - orc_S).max(1)
    allow = FP32_SITE_TOL + SAFETY * E
    ratio = np.where(taint, 0.0, diff / allow)
    worst = int(np.argmax(ratio))
    print(f"fp32 bound {label}: {seeds} possible flips in {n_updates} updates, tainted {int(taint.sum())} of "
          f"{len(E)}, sites off by > 1e-3: {int((diff > 1e-3).sum())}, max dev {diff.max():.3g} (untainted "
          f"{diff[~taint].max() if (~taint).any() else 0:.3g}), max dev/allowance {ratio[worst]:.3g} "
          f"(site {worst}, bound {E[worst]:.3g})")
    assert seeds <= max(4, n_updates // 100), seeds
    assert ratio[worst] <= 1.0, (worst, diff[worst], E[worst])
    return float(ratio[worst])
