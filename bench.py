#!/usr/bin/env python3
"""Headline benchmark: batched FP32 parallel tempering of the pyrochlore
J1-J2-J3 Heisenberg model on one MI355X.

Workload (the reference's tempering sweep, temper.cpp:418-474, with all
replicas on one GPU): L = 12 (27,648 spins), 64 replicas at log-spaced T in
[0.02, 2.0], J1 = 1, J2 = -0.05, J3 = 0.05, B = 0; each sweep is one
Gaussian-kick Metropolis pass + two overrelaxation passes; one step = one
tempering round = 4 sweeps, one SSF sample of all 9 correlators at every
rung, the energy sample and the neighbour-pair exchange.
Metric: site-updates/s = N * R * 3 passes * 4 sweeps * steps / wall time.

Also reported: the plain FP64 one-replica GPU path on the same lattice and
sweep, the sweep kernel's HBM roofline, the sweep kernel of the large
tempering lattice (L = 48 x 16 replicas, lifted + OR), the sublattice-resolved
correlator products on FP32 MFMA at that size, and the CPU restatement of the
reference sweep timed on this host.

HBM traffic per launch ("traffic") comes from the rocprofv3 FETCH_SIZE and
WRITE_SIZE summaries committed under profiles/ for the same kernel and
workload (lines "traffic tag=<tag> bytes_per_launch=<B>"), or is null.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs c3,fp64,large,munu,c5,batch,cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import re

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

L_LAT = 12
R_REP = 64
J123 = (1.0, -0.05, 0.05)
T_RANGE = (2.0, 0.02)
N_OR = 2
N_SWEEP = 4
CORR = ("xx", "yy", "zz", "xy", "xz", "yz", "yx", "zx", "zy")
PEAK_HBM = 8.0e12  # B/s, MI355X HBM3E spec
PEAK_FP32_MATRIX = 157.3e12  # flop/s, MI355X dense FP32 MFMA spec
ALL_LEGS = ("c3", "fp64", "large", "munu", "c5", "batch", "cpu")


def profiled_traffic(tag: str):
    """HBM bytes per launch measured for `tag` by the rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs summarised in profiles/*.txt, or None."""
    pat = re.compile(r"traffic tag=(\S+) bytes_per_launch=([0-9.eE+]+)")
    for f in sorted((ROOT / "profiles").glob("*.txt")):
        for line in f.read_text().splitlines():
            m = pat.search(line)
            if m and m.group(1) == tag:
                return float(m.group(2))
    return None


def bytes_per_update(prec_bytes: int, n_nbr: int, r_share: int, lifted_frac: float = 0.0) -> float:
    """Compulsory HBM bytes per site-update: read + write own spin (3
    components), the int32 neighbour row amortised over the replicas that
    read it in one launch, and the lifted direction byte (read + write)."""
    return 2 * 3 * prec_bytes + 4.0 * n_nbr / r_share + 2.0 * lifted_frac


# The tempering workload's sweeps run in the fused pass order
# (pyro_set_fused_passes): per colour, each site's Metropolis update and its
# first OR update in one launch from one field evaluation, then the second
# OR pass on its own: 32 colour launches per sweep instead of 48.  The FP64
# one-replica path it is compared with keeps the reference's pass-by-pass
# order.
C3_FUSED = True
C3_ORDER = ("fused: per colour, each site's Metropolis and first OR update from one field evaluation, "
            "then the second OR pass" if C3_FUSED else "pass by pass")


def make_c3(pyro, precision, R, seed=2024, fused=False):
    ctx = pyro.Context.pyrochlore(L_LAT, n_replicas=R, precision=precision, seed=seed)
    ctx.define_J1J2J3(*J123)
    ctx.setup()
    ctx.set_fused_passes(fused)
    ctx.randomize(seed + 1)
    ladder = np.geomspace(T_RANGE[0], T_RANGE[1], R) if R > 1 else np.array([T_RANGE[1]])
    ctx.pt_init(ladder, swap_seed=seed + 2)
    return ctx


def time_rounds(ctx, pyro, steps, warmup, sample_ssf=True):
    ctx.pt_rounds(pyro.METROPOLIS, N_OR, N_SWEEP, sample_ssf, warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.mark(0)
    ctx.pt_rounds(pyro.METROPOLIS, N_OR, N_SWEEP, sample_ssf, steps)
    ctx.mark(1)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    return wall, ctx.elapsed_ms() / 1e3


def large_lattice_sweep(pyro, sweeps=20, reps=3):
    """Sweep kernel on the large tempering lattice: L = 48 (1,769,472 spins)
    x 16 replicas, FP32, J1 = 1, J2 = J3 = 0.1, B = (0, 0, 0.2), T log-spaced
    in [0.05, 1], each sweep 1 lifted Metropolis + 1 overrelaxation pass, no
    acceptance counting (as in the tempering rounds); the best of `reps`
    stretches of `sweeps` sweeps, HIP events around each.  Returns
    (site-updates/s, per-launch HIP-event timing, strategy)."""
    R = 16
    ctx = pyro.Context.pyrochlore(48, n_replicas=R, precision=pyro.FP32, seed=48)
    ctx.define_J1J2J3(1.0, 0.1, 0.1)
    ctx.setup()
    ctx.set_fused_passes(LARGE_FUSED)
    ctx.set_field(np.array([0.0, 0.0, 0.2]))
    ctx.randomize(49)
    ctx.set_temperatures(np.geomspace(1.0, 0.05, R))
    ctx.sweep(pyro.LIFTED, 1, 2, count=False)
    ctx.synchronize()
    ms = float("inf")
    for _ in range(reps):
        ctx.mark(0)
        ctx.sweep(pyro.LIFTED, 1, sweeps, count=False)
        ctx.mark(1)
        ms = min(ms, ctx.elapsed_ms())
    updates = ctx.N * R * 2 * sweeps
    ctx.profile(True)
    ctx.sweep(pyro.LIFTED, 1, 2, count=False)
    prof = ctx.profile_read()
    ctx.profile(False)
    sweep_ms, launches = prof["sweep"]
    strategy = ctx.strategy()
    n = ctx.N
    ctx.close()
    upd_per_launch = n * R * 2 * 2 / launches
    return updates / (ms / 1e3), sweep_ms / launches, upd_per_launch, strategy


def munu_correlators(pyro, samples=384, buffer=192):
    """Sublattice-resolved correlators of the large tempering lattice:
    L = 48 x 16 replicas, FP32 spins, all 9 (alpha, beta) pairs over the 16
    sublattices (the 48 x 48 Hermitian matrix per k), one accumulator slot
    per replica.  Samples are staged as FP32 rows X = [Re | Im] of the 48
    transformed fields, `buffer` deep, and flushed as X^T X on FP32 MFMA
    into FP64 sums (tile pairs I <= J of the Hermitian matrix), for the
    kept half k2 <= L/2 of the K (M(-K) = conj(M(K)) for real spins).  192
    rows per flush: 68 GB of staging beside 22.6 GB of FP64 accumulators, so
    the accumulator round trip is spread over three times as many samples as
    at 64.  Returns the flush kernel's timing and work."""
    R, L = 16, 48
    ctx = pyro.Context.pyrochlore(L, n_replicas=R, precision=pyro.FP32, seed=48)
    ctx.define_J1J2J3(1.0, 0.1, 0.1)
    ctx.setup()
    ctx.randomize(49)
    ctx.ssf_configure(CORR, n_slots=R, store_error_term=True, munu=True, munu_buffer=buffer)
    ctx.profile(True)
    for _ in range(samples):
        ctx.ssf_sample()
    ctx.synchronize()
    prof = ctx.profile_read()
    ctx.profile(False)
    ctx.close()
    flush_ms, n_flush = prof["munu_flush"]
    n_k, T = L * L * (L // 2 + 1), 3  # the kept K; 48 features = 3 tiles of 16
    pairs = T * (T + 1) // 2  # tile pairs I <= J of the Hermitian matrix
    # per row and tile pair: RR, II, RI, IR products, 16 x 16 x 2 flops each
    flops = R * n_k * buffer * pairs * 4 * 16 * 16 * 2
    acc_bytes = R * n_k * pairs * 256 * 16 * 2  # FP64 complex accumulators read + written
    row_bytes = R * n_k * buffer * 2 * 16 * T * 4  # staged FP32 rows read
    return {
        "workload": f"L={L} (1769472 spins) x {R} replicas, FP32, 9 correlators over 16 sublattices, "
                    f"one slot per replica, {buffer} samples buffered per flush",
        "flush_avg_ms": flush_ms / max(n_flush, 1),
        "flushes": n_flush,
        # FP32: the staging runs inside the contraction kernel (one read of the fields)
        "contract_and_stage_ms_per_sample": (prof["contract"][0] + prof["munu"][0]) / max(samples, 1),
        "dft_ms_per_sample": prof["dft"][0] / max(samples, 1),
        "roofline": {
            "bound": "mfma",
            "achieved": flops / (flush_ms / max(n_flush, 1) / 1e3) / 1e12,
            "peak": PEAK_FP32_MATRIX / 1e12,
            "unit": "TFLOP/s",
            "frac": flops / (flush_ms / max(n_flush, 1) / 1e3) / PEAK_FP32_MATRIX,
            "flops_per_launch": flops,
            "bytes_per_launch": acc_bytes + row_bytes,
            "traffic": profiled_traffic("munu_flush_L48x16"),
        },
    }


# The large runs' sweeps take a site's lifted and OR updates back to back
# within its colour stage (pyro_set_fused_passes): one field evaluation and
# one Philox block per site and sweep instead of two.
LARGE_FUSED = True
LARGE_ORDER = ("fused: per tile and colour, each site's lifted then OR update from one field evaluation"
               if LARGE_FUSED else "pass by pass within each tile")


def large_tempering_run(pyro, rounds=192, buffer=192):
    """The large tempering workload end to end: L = 48 (1,769,472 spins) x 16
    replicas, FP32, J1 = 1, J2 = J3 = 0.1, B = (0, 0, 0.2), T log-spaced in
    [0.05, 1]; each round is two sweeps (1 lifted Metropolis + 1 OR pass
    each) with an SSF sample of all 9 correlators over the 16 sublattices
    after each sweep (contracted sums and the sublattice-resolved matrices,
    buffered `buffer` deep and flushed on MFMA into FP64), then the
    energies (one sample per round, as temper.cpp:432-435 takes one per
    exchange) and a replica exchange; captured as one graph per round.  The
    timed region holds `rounds` rounds and the flushes they trigger."""
    R, L = 16, 48
    ctx = pyro.Context.pyrochlore(L, n_replicas=R, precision=pyro.FP32, seed=48)
    ctx.define_J1J2J3(1.0, 0.1, 0.1)
    ctx.setup()
    ctx.set_fused_passes(LARGE_FUSED)
    ctx.set_field(np.array([0.0, 0.0, 0.2]))
    ctx.randomize(49)
    ctx.pt_init(np.geomspace(1.0, 0.05, R), 50)
    ctx.ssf_configure(CORR, n_slots=R, store_error_term=True, munu=True, munu_buffer=buffer)
    ctx.pt_rounds(pyro.LIFTED, 1, 2, 2, 2)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.pt_rounds(pyro.LIFTED, 1, 2, 2, rounds)  # SSF after each of the 2 sweeps
    ctx.synchronize()
    wall = time.perf_counter() - t0
    n = ctx.N
    st = ctx.pt_stats()
    # per-kernel HIP-event times of 8 rounds run eagerly (no flush among them)
    ctx.profile(True)
    ctx.pt_rounds(pyro.LIFTED, 1, 2, 2, 8)
    prof = ctx.profile_read()
    ctx.profile(False)
    ctx.close()
    sweeps = 2 * rounds
    return {
        "workload": f"parallel tempering, L={L} (1769472 spins) x {R} replicas, FP32, J1=1 J2=J3=0.1, "
                    "B=(0,0,0.2), T log-spaced 1.0..0.05, sweep = 1 lifted Metropolis + 1 OR, SSF of 9 "
                    f"correlators over 16 sublattices every sweep ({buffer}-sample MFMA flushes into FP64), "
                    "exchange and energy sample every 2 sweeps",
        "pass_order": LARGE_ORDER,
        "site_updates_per_s": n * R * 2 * sweeps / wall,
        "ms_per_sweep": 1e3 * wall / sweeps,
        "sweeps": sweeps,
        "swap_acceptance_min": float(np.min(st["accepts"] / np.maximum(st["attempts"], 1))),
        "kernel_ms_per_sweep": {k: v[0] / 16 for k, v in prof.items() if v[1]},
    }


# The uneven batch's Metropolis + OR sweeps in the fused pass order: one
# stage per colour (a launch per colour on the small lattices, a colour
# stage per tile on L = 16 and 24) instead of two.
BATCH_FUSED = True


def uneven_batch(pyro):
    """256 independent anneals with L drawn from {4, 8, 12, 16, 24}, J1 = 1,
    J2, J3 uniform in [-0.2, 0.2], |B| uniform in [0, 0.5] along a random
    direction; T 10 -> 0.05 over 100 log steps, 16 burn-in sweeps, 16 sweeps
    per step, 64 SSF samples (xx, yy, zz) at T_cold; FP32 spins, FP64 sums;
    Metropolis + OR.  Items of equal L share a context; the five contexts
    step together as one Batch (one graph per step: L = 4, 8, 12 in one
    persistent launch, L = 16 and 24 tiled beside it).  The same anneals with
    each context replaying its own graph on its own stream are timed after
    it for comparison."""
    from pyro_cmc_amd import batch

    items = batch.uneven_items(256, seed=4)
    _, st = batch.run_anneal_batch(items, seed=4, fused=BATCH_FUSED)
    _, own = batch.run_anneal_batch(items, seed=4, fused=BATCH_FUSED, persistent=False)
    return {
        "workload": "256 anneals, L in {4,8,12,16,24}, random J2/J3 in [-0.2,0.2], |B| in [0,0.5], "
                    "T 10->0.05 over 100 steps, 16 sweeps/step, 64 SSF samples at T_cold, FP32",
        "pass_order": "fused: per colour, each site's Metropolis and OR update from one field evaluation"
                      if BATCH_FUSED else "pass by pass",
        "site_updates_per_s": st["site_updates_per_s"],
        "wall_s": st["wall_s"],
        "site_updates": st["site_updates"],
        "items_per_L": st["contexts"],
        "persistent_launch": st.get("persistent_launch"),
        "own_graphs_site_updates_per_s": own["site_updates_per_s"],
    }


def _cpu_worker(job):
    """One host process: the oracle's restatement of the reference sweep on
    its own L = 12 replica for `seconds`; returns (site updates, seconds)."""
    seconds, seed = job
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle

    o = pyoracle.Oracle.pyrochlore(L_LAT, *J123, seed=seed)
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(o.N, 3))
    o.set_spins(v / np.linalg.norm(v, axis=1, keepdims=True))
    o.sweep_ref(0, 0.5, N_OR)  # warm-up
    n = 0
    t0 = time.perf_counter()
    while True:
        o.sweep_ref(0, 0.5, N_OR)
        n += 1
        dt = time.perf_counter() - t0
        if dt > seconds:
            break
    return o.N * (1 + N_OR) * n, dt


def _cpu_worker_best(job):
    """One host process alone: warm-up, then `reps` timed stretches of
    `seconds` each of the same sweeps; returns the best rate
    (bench_speed.cpp:173-187 reports the best of its repetitions)."""
    seconds, seed, reps = job
    rates = []
    for k in range(reps):
        u, dt = _cpu_worker((seconds, seed + 1000 * k))
        rates.append(u / dt)
    return max(rates), rates


def cpu_baseline(seconds: float = 12.0):
    """The oracle's restatement of the reference sweep (sequential, AoS
    spins with bond-shell pointer lists, Xoroshiro128++ + libstdc++
    distributions; MC.cpp:87-163), one L = 12 replica per host process,
    Metropolis + 2 OR per sweep at T = 0.5, as the reference runs one
    replica per process.  Two measurements: one process alone, best of 3
    stretches of seconds / 3 (the rate of the reference on one core), then
    as many processes as cores (up to 16) for `seconds` at once (the rate of
    the reference run as independent seed jobs, acc_runs.py:13-57), whose
    aggregate is `value`.  Worker processes are started fresh (spawn),
    never forked from this GPU process."""
    import multiprocessing as mp

    ctx = mp.get_context("spawn")
    with ctx.Pool(1) as pool:
        single, single_reps = pool.map(_cpu_worker_best, [(seconds / 3.0, 5, 3)])[0]
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    with ctx.Pool(workers) as pool:
        res = pool.map(_cpu_worker, [(seconds, 7 + i) for i in range(workers)])
    total = sum(u for u, _ in res) / max(dt for _, dt in res)
    return {
        "value": total,
        "unit": "site-updates/s",
        "cores": workers,
        "kind": "port",
        "single_core_value": single,
        "single_core_reps": single_reps,
        "sample": f"{workers} processes at once, each one L={L_LAT} replica ({16 * L_LAT ** 3} spins) swept with "
                  f"Metropolis + {N_OR} OR at T=0.5 for {seconds:.0f} s (value); one process alone, best of 3 "
                  f"stretches of {seconds / 3:.0f} s (single_core_value); host {os.cpu_count()} logical CPUs, "
                  f"{len(os.sched_getaffinity(0))} available",
    }


def large_lattice_sweep_line(pyro):
    rate, launch_ms, upl, strat_big = large_lattice_sweep(pyro)
    # 1 lifted + 1 OR pass: the lifted direction byte (read + write) on half
    # of the passes; the int32 table term is zero for the stencil kernels (no
    # table is read), kept at its formula value for the comparison with the
    # table kernels
    b_big = bytes_per_update(4, 30, 16, lifted_frac=0.5)
    traffic = profiled_traffic("sweep_L48x16_lifted")
    return {
        "workload": "L=48 (1769472 spins) x 16 replicas, FP32, J1=1 J2=J3=0.1, B=(0,0,0.2), "
                    "sweep = 1 lifted Metropolis + 1 OR",
        "pass_order": LARGE_ORDER,
        "site_updates_per_s": rate,
        "strategy": strat_big,
        "roofline": {
            "bound": "hbm", "achieved": b_big * upl / (launch_ms / 1e3) / 1e9,
            "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": b_big * upl / (launch_ms / 1e3) / PEAK_HBM,
            # a tiled launch reads and writes each spin (and its direction
            # byte) once for both passes
            "frac_per_site_visit": bytes_per_update(4, 30, 16, lifted_frac=1.0) * (upl / 2)
                                   / (launch_ms / 1e3) / PEAK_HBM,
            "traffic": traffic,
            # the measured bytes per launch moved in the measured launch time
            "traffic_rate_frac": (traffic / (launch_ms / 1e3) / PEAK_HBM) if traffic else None,
            "bytes_per_update": b_big, "updates_per_launch": upl,
            "avg_launch_us": launch_ms * 1e3,
        },
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1024)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the L=48 sweep-kernel line")
    ap.add_argument("--no-munu", action="store_true", help="skip the L=48 MFMA correlator line")
    ap.add_argument("--legs", default=",".join(ALL_LEGS),
                    help="comma list of parts to run: " + ",".join(ALL_LEGS))
    ap.add_argument("--strategy", type=int, default=0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (counter calibration)")
    args = ap.parse_args()

    if args.copy_probe:
        import torch

        x = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 << 28}))
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the tempering "
              "path runs all replicas on a single device.", file=sys.stderr)
        sys.exit(2)

    import pyro_cmc_amd as pyro

    if pyro.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)

    legs = set(args.legs.split(","))
    for flag, leg in ((args.no_cpu, "cpu"), (args.no_fp64, "fp64"), (args.no_large, "large"),
                      (args.no_munu, "munu")):
        if flag:
            legs.discard(leg)
    if "c3" not in legs:  # profiling runs of a single leg
        out = {}
        if "large" in legs:
            out["large_lattice_sweep"] = large_lattice_sweep_line(pyro)
        if "munu" in legs:
            out["munu_correlators"] = munu_correlators(pyro)
        if "c5" in legs:
            out["large_tempering_run"] = large_tempering_run(pyro)
        if "batch" in legs:
            out["uneven_batch"] = uneven_batch(pyro)
        print(json.dumps(out))
        return

    # ---- headline: batched FP32 tempering ----
    ctx = make_c3(pyro, pyro.FP32, R_REP, fused=C3_FUSED)
    ctx.ssf_configure(CORR, n_slots=R_REP, store_error_term=True)
    ctx.set_strategy(args.strategy)
    wall, gpu_s = time_rounds(ctx, pyro, args.steps, args.warmup)
    N = ctx.N
    updates = N * R_REP * (1 + N_OR) * N_SWEEP * args.steps
    value = updates / wall
    strategy = ctx.strategy()
    info = ctx.info()

    # ---- per-kernel timing with HIP events on the context stream ----
    prof_steps = max(8, min(args.steps, 256))
    ctx.profile(True)
    ctx.pt_rounds(pyro.METROPOLIS, N_OR, N_SWEEP, True, prof_steps)
    prof = ctx.profile_read()
    ctx.profile(False)
    # Sweep-kernel launch duration as the timed rounds run it: the profiling
    # mode's HIP events around each round's sweep launches (the energy
    # variant, after the previous round's SSF, with the XCD-resident sweep's
    # counter clear included), on the context stream.
    sweep_prof_ms, sweep_prof_launches = prof["sweep"]
    launches_per_round = sweep_prof_launches / prof_steps
    avg_launch_s = sweep_prof_ms / 1e3 / sweep_prof_launches
    # colour stages per sweep: one per colour and pass, or in the fused order
    # one per colour for the Metropolis + first OR update and one per colour
    # for each further OR pass
    groups = max(N_OR, 1) if C3_FUSED else 1 + N_OR
    upd_per_launch = N * R_REP * (1 + N_OR) * N_SWEEP / launches_per_round
    sites_per_launch = N * R_REP * N_SWEEP * (groups if strategy in (3, 6) else 1 + N_OR) / launches_per_round
    # Beside it: back-to-back launches of a round's sweeps with no per-launch
    # events (the plain variant, warm caches), as a second sample.
    n_sw = max(16, min(args.steps, 256))
    ctx.synchronize()
    ctx.mark(0)
    if strategy == 6:
        for _ in range(n_sw // N_SWEEP):
            ctx.sweep(pyro.METROPOLIS, N_OR, N_SWEEP, count=False)
    else:
        ctx.sweep(pyro.METROPOLIS, N_OR, n_sw, count=False)
    ctx.mark(1)
    sweep_ms = ctx.elapsed_ms()
    b2b_launches = {1: n_sw * (1 + N_OR) * info["n_colours"], 3: n_sw * groups * info["n_colours"],
                    2: 1, 4: 1, 5: 8 * n_sw, 6: n_sw // N_SWEEP}[strategy]
    # compulsory bytes per site-update (its spin read and written, its share
    # of the neighbour row) x the launch's site updates; beside it the
    # stricter count per site visited: a fused stage reads and writes a spin
    # once for its two updates
    b_alg = bytes_per_update(4, info["n_slot"], R_REP)
    achieved = b_alg * upd_per_launch / avg_launch_s
    achieved_visit = b_alg * sites_per_launch / avg_launch_s
    traffic_c3 = profiled_traffic("sweep_c3_L12x64" + ("_xcd" if strategy == 6 else ""))
    st = ctx.pt_stats()
    acc_rate = (st["accepts"] / np.maximum(st["attempts"], 1)).tolist()

    # ---- FP64 one-replica GPU path, same lattice and sweep ----
    fp64 = None
    if "fp64" in legs:
        c1 = make_c3(pyro, pyro.FP64, 1)
        steps64 = max(4, args.steps // 16)
        w64, _ = time_rounds(c1, pyro, steps64, 4, sample_ssf=False)
        fp64 = c1.N * (1 + N_OR) * N_SWEEP * steps64 / w64
        c1.close()

    cpu = cpu_baseline(args.cpu_seconds) if "cpu" in legs else None
    big = large_lattice_sweep_line(pyro) if "large" in legs else None
    munu = munu_correlators(pyro) if "munu" in legs else None
    c5 = large_tempering_run(pyro) if "c5" in legs else None
    unev = uneven_batch(pyro) if "batch" in legs else None

    line = {
        "metric": "site-updates/s (Metropolis+OR) at L=12x64 replicas",
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: spins uniform on the sphere from the seed",
        "config": {
            "workload": "parallel tempering, pyrochlore L=12 (27648 spins) x 64 replicas, "
                        "T log-spaced 2.0..0.02, J1=1 J2=-0.05 J3=0.05, sweep = 1 Metropolis + "
                        "2 OR, exchange every 4 sweeps, SSF (9 correlators) every round",
            "L": L_LAT, "replicas": R_REP, "passes_per_sweep": 1 + N_OR,
            "sweeps_per_step": N_SWEEP, "strategy": strategy, "pass_order": C3_ORDER,
        },
        "gpu_event_s": gpu_s,
        "fp64_one_replica_site_updates_per_s": fp64,
        "speedup_vs_fp64_one_replica": (value / fp64) if fp64 else None,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": PEAK_HBM / 1e9,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM,
            "frac_per_site_visit": achieved_visit / PEAK_HBM,
            "traffic": traffic_c3,
            # the measured bytes per launch moved in the measured launch time
            "traffic_rate_frac": (traffic_c3 / avg_launch_s / PEAK_HBM) if traffic_c3 else None,
            "kernel": {1: "sweep_colour_kernel", 2: "sweep_resident_kernel", 3: "pyro_colour_kernel",
                       4: "pyro_resident_kernel", 5: "pyro_tiled_kernel", 6: "pyro_xcd_kernel"}[strategy],
            "bytes_per_update": b_alg,
            "updates_per_launch": upd_per_launch,
            "sites_visited_per_launch": sites_per_launch,
            "avg_launch_us": avg_launch_s * 1e6,
            # the HIP-event sample avg_launch_us comes from
            "launch_timing": {"rounds": prof_steps, "launches": sweep_prof_launches, "elapsed_ms": sweep_prof_ms,
                              "launches_per_round": launches_per_round},
            # back-to-back launches of the plain variant, no per-launch events
            "back_to_back": {"sweeps": n_sw, "launches": b2b_launches, "elapsed_ms": sweep_ms,
                             "avg_launch_us": 1e3 * sweep_ms / b2b_launches},
        },
        "kernel_ms": {k: v[0] for k, v in prof.items()},
        "kernel_profiled_rounds": prof_steps,
        "kernel_launches": {k: v[1] for k, v in prof.items()},
        # 64 log-spaced rungs over two decades at 27,648 spins: adjacent rungs'
        # energy distributions barely overlap, so swaps are rare (about 1%
        # at the hot end, none at the cold end)
        "swap_acceptance_min": float(min(acc_rate)) if acc_rate else None,
        "swap_acceptance_max": float(max(acc_rate)) if acc_rate else None,
        "swap_attempts_per_pair": float(np.mean(st["attempts"])) if acc_rate else None,
        "large_lattice_sweep": big,
        "munu_correlators": munu,
        "large_tempering_run": c5,
        "uneven_batch": unev,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
