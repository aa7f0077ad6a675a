"""Uneven batches of independent anneals on one GPU.

The reference runs one anneal per process (anneal.cpp:33-248) and merges
seeds afterwards (acc_runs.py:13-57).  Here a batch of models with different
lattice sizes, couplings and fields runs as one job: items of equal L share a
context (one replica per item, per-replica couplings and field), and every
context follows the same device-side temperature schedule.  The contexts
step together as a Batch: each anneal step of the whole batch is one
replayed HIP graph, in which the contexts on the per-colour stencil
strategies (the small lattices) sweep in one persistent launch and the
large lattices run their tiled sweeps beside it.

Schedule per item, as in anneal.cpp:193-221: n_burn_in sweeps at T_hot;
then for every T of generate_T_profile(T_hot, T_cold, T_sample, n_steps)
n_sweep sweeps and one energy sample, and where T is one of the T_sample
entries n_sample x (n_sweep sweeps + one SSF sample) at T into that
temperature's accumulator (ssf_manager::new_T).  T_sample defaults to
{T_cold}.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import FP32, METROPOLIS, Batch, Context


@dataclass
class AnnealItem:
    L: int
    J: tuple  # (J1, J2, J3, Delta)
    B: tuple = (0.0, 0.0, 0.0)


@dataclass
class AnnealResult:
    item: AnnealItem
    T_list: np.ndarray              # the anneal grid, descending
    E: np.ndarray                   # energy per cell after each grid step
    T_sample: np.ndarray            # sampling temperatures, descending (ssf_manager T order)
    S: np.ndarray                   # SSF sums per T_sample, [n_T][n_corr][n_k][2]
    S2: np.ndarray                  # (Re S)^2, (Im S)^2 sums, same shape
    n_samples: np.ndarray           # samples per T_sample, [n_T]
    E_final: float                  # energy per cell of the final configuration
    spins: np.ndarray | None = field(default=None, repr=False)  # final [N][3] when kept


def generate_T_profile(T_hot, T_cold, T_sample, total_steps):
    """Log grid T_hot * f^i, i < total_steps, f = (T_cold/T_hot)^(1/total_steps),
    merged in descending order with T_sample (cli_bits.hpp:250-289); returns
    the grid and the indices of the T_sample entries."""
    f = (T_cold / T_hot) ** (1.0 / total_steps)
    sweep = sorted((T_hot * f ** i for i in range(total_steps)), reverse=True)
    samp = sorted(T_sample, reverse=True)
    out, idx, i, j = [], [], 0, 0
    while i < len(sweep) and j < len(samp):
        if sweep[i] > samp[j]:
            out.append(sweep[i])
            i += 1
        else:
            idx.append(len(out))
            out.append(samp[j])
            j += 1
    for t in samp[j:]:
        idx.append(len(out))
        out.append(t)
    out.extend(sweep[i:])
    return np.array(out), idx


def uneven_items(n=256, sizes=(4, 8, 12, 16, 24), seed=0):
    """The uneven batch: L uniform over `sizes`, J1 = 1, J2 and J3 uniform in
    [-0.2, 0.2], |B| uniform in [0, 0.5] along a uniform random direction."""
    rng = np.random.default_rng(seed)
    items = []
    for _ in range(n):
        L = int(rng.choice(sizes))
        J2, J3 = rng.uniform(-0.2, 0.2, 2)
        v = rng.normal(size=3)
        B = v / np.linalg.norm(v) * rng.uniform(0.0, 0.5)
        items.append(AnnealItem(L, (1.0, float(J2), float(J3), 1.0), tuple(float(b) for b in B)))
    return items


def run_anneal_batch(items, T_hot=10.0, T_cold=0.05, n_steps=100, n_sweep=16, n_burn_in=16, n_sample=64,
                     n_overrelax=1, kind=METROPOLIS, precision=FP32, correlators=("xx", "yy", "zz"),
                     seed=0, keep_spins=False, strategy=0, T_sample=None, fused=False, persistent=True):
    """Run every item's anneal; returns (results in item order, stats).
    fused: the sweeps in the fused pass order (pyro_set_fused_passes: per
    colour stage each site's first-kind and first OR update from one field
    evaluation) instead of pass by pass.  persistent: the contexts step
    together as a Batch (one graph per step, the small lattices sweeping in
    one persistent launch) instead of each replaying its own graph; the
    results are the same bit for bit."""
    T_sample = [T_cold] if T_sample is None else list(T_sample)
    grid, s_idx = generate_T_profile(T_hot, T_cold, T_sample, n_steps)
    n_ts = len(s_idx)
    # device schedule: the burn-in step at T_hot, then every grid step, each
    # sampling step followed by its n_sample SSF steps at the same T
    sched = [T_hot]
    e_at = []  # schedule index of each grid step's energy sample
    phases = [(n_burn_in, False, 1, -1)]  # (sweeps, ssf, steps, T_sample index)
    run = 0
    for i, T in enumerate(grid):
        e_at.append(len(sched))
        sched.append(T)
        run += 1
        if i in s_idx:
            phases.append((n_sweep, False, run, -1))
            run = 0
            phases.append((n_sweep, True, n_sample, s_idx.index(i)))
            sched.extend([T] * n_sample)
    if run:
        phases.append((n_sweep, False, run, -1))
    sched = np.array(sched)
    groups = {}
    for i, it in enumerate(items):
        groups.setdefault(it.L, []).append(i)
    ctxs = []
    for L, idx in sorted(groups.items()):
        c = Context.pyrochlore(L, n_replicas=len(idx), precision=precision, seed=seed * 1000003 + L)
        # the context's shells must cover every item's: all three, with zeros allowed per replica
        c.define_J1J2J3(1.0, 1.0, 1.0)
        c.setup()
        c.set_strategy(strategy)
        c.set_fused_passes(fused)
        c.set_replica_couplings(np.array([items[i].J for i in idx], np.float64))
        for r, i in enumerate(idx):
            c.set_field(np.array(items[i].B, np.float64), r)
        c.randomize(seed * 7919 + L)
        # slot j R + r holds replica r's samples at T_sample j
        c.ssf_configure(correlators, n_slots=n_ts * len(idx), store_error_term=True)
        c.anneal_schedule(sched)
        ctxs.append((c, idx))
    batch = Batch([c for c, _ in ctxs]) if persistent else None
    for c, _ in ctxs:
        c.synchronize()
    # timed: everything from the first step on, including each context's
    # strategy choice and graph capture on its first call
    t0 = time.perf_counter()
    for n_sw, ssf, steps, j in phases:
        for c, _ in ctxs:
            if ssf and n_ts > 1:  # (setting a map waits for the context's queued steps)
                c.pt_ssf_map(np.arange(c.R, dtype=np.int32) + j * c.R)
            if batch is None:  # enqueue only: the contexts' streams run side by side
                c.anneal_steps(kind, n_overrelax, n_sw, ssf, steps)
        if batch is not None:
            batch.anneal_steps(kind, n_overrelax, n_sw, ssf, steps)
    for c, _ in ctxs:
        c.synchronize()
    wall = time.perf_counter() - t0
    launch = batch.info() if batch is not None else None
    if batch is not None:
        batch.close()
    results = [None] * len(items)
    updates = 0
    e_at = np.array(e_at)
    for c, idx in ctxs:
        hist = c.anneal_history()
        for r, i in enumerate(idx):
            reads = [c.ssf_read(j * c.R + r) for j in range(n_ts)]
            results[i] = AnnealResult(items[i], grid.copy(), hist[r, e_at].copy(), grid[s_idx].copy(),
                                      np.array([q[0] for q in reads]), np.array([q[1] for q in reads]),
                                      np.array([q[2] for q in reads]), float(hist[r, -1]),
                                      c.get_spins(r) if keep_spins else None)
        passes = (1 + n_overrelax) * (n_burn_in + n_sweep * (len(sched) - 1))
        updates += c.N * c.R * passes
        c.close()
    stats = {"wall_s": wall, "site_updates": updates, "site_updates_per_s": updates / wall,
             "contexts": {int(L): len(idx) for L, idx in sorted(groups.items())}, "schedule_steps": len(sched)}
    if launch is not None:
        stats["persistent_launch"] = {"contexts": launch[0], "blocks_per_workgroup": launch[1],
                                      "workgroups": launch[2]}
    return results, stats
