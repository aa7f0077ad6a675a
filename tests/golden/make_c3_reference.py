"""Reference-sampler statistics at three rungs of the L = 12 tempering
workload (J = (1, -0.05, 0.05), Metropolis + 2 OR per sweep, T_ref = 1), for
the GPU comparison in tests/test_workloads.py.

Each rung is sampled by 16 independent single-temperature chains of the CPU
oracle in reference mode: storage-order sweeps with Xoroshiro128++ driving
the libstdc++ normal / uniform distributions, as MC.cpp:155-163 sweeps,
started from uniform spins.  Per chain: N_BURN sweeps, then N_MEAS sweeps
with the energy and S(q) taken every 4th sweep (the workload's rounds are 4
sweeps).  Per chain it records <E>/N, C_v/N = N var(E/N) / T^2 and S(q)/N =
sum_a |sum_i e^{-iq.r_i} S^a_i|^2 / N at K = (0,0,0), (0,0,1), (0,1,1), with
q(K) = 2 pi K' / (8 L) on the centred index; the summary is the mean over
chains and its seed-to-seed standard error (acc_runs.py:315-331).

Run from the repository root:  python tests/golden/make_c3_reference.py
(about 5 minutes on 8 cores); writes tests/golden/c3_reference.json.
"""
from __future__ import annotations

import json
import sys
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
L, R_LADDER = 12, 64
J = (1.0, -0.05, 0.05)
RUNGS = (0, 16, 32)  # T = 2.0, 0.621, 0.193 of geomspace(2, 0.02, 64)
N_SEEDS = 16
N_BURN, N_MEAS, EVERY = 2000, 4000, 4
KS = ((0, 0, 0), (0, 0, 1), (0, 1, 1))


def chain(args):
    rung, seed = args
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle

    T = float(np.geomspace(2.0, 0.02, R_LADDER)[rung])
    o = pyoracle.Oracle.pyrochlore(L, *J, seed=1000 * rung + seed)
    o.randomize_ctr(777 + seed, rung)
    pos = o.positions().astype(float)
    q = []
    for K in KS:
        Kc = [k if 2 * k < L else k - L for k in K]
        q.append(2 * np.pi * np.array(Kc) / (8.0 * L))
    ph = np.exp(-1j * pos @ np.array(q).T)  # [N][3 K]
    for _ in range(N_BURN):
        o.sweep_ref(0, T, 2)
    e, s = [], []
    for k in range(N_MEAS):
        o.sweep_ref(0, T, 2)
        if (k + 1) % EVERY == 0:
            e.append(o.energy_per_cell() / 16.0)  # per spin
            S = o.get_spins()
            b = ph.T @ S  # [K][a]
            s.append((np.abs(b) ** 2).sum(1) / o.N)
    e = np.array(e)
    s = np.array(s)
    return rung, seed, T, float(e.mean()), float(o.N * e.var() / T ** 2), [float(x) for x in s.mean(0)]


def main():
    jobs = [(r, s) for r in RUNGS for s in range(N_SEEDS)]
    with ProcessPoolExecutor(8) as ex:
        out = list(ex.map(chain, jobs))
    res = {"L": L, "J": J, "ladder": "geomspace(2.0, 0.02, 64)", "kind": "Metropolis + 2 OR", "n_seeds": N_SEEDS,
           "n_burn": N_BURN, "n_meas": N_MEAS, "every": EVERY, "K": KS, "rungs": []}
    for r in RUNGS:
        rows = [x for x in out if x[0] == r]
        T = rows[0][2]
        e = np.array([x[3] for x in rows])
        cv = np.array([x[4] for x in rows])
        s = np.array([x[5] for x in rows])
        n = len(rows)
        res["rungs"].append({
            "rung": r, "T": T,
            "e_per_spin": [float(e.mean()), float(e.std(ddof=1) / np.sqrt(n))],
            "cv_per_spin": [float(cv.mean()), float(cv.std(ddof=1) / np.sqrt(n))],
            "S_per_N": [[float(m), float(d)] for m, d in zip(s.mean(0), s.std(0, ddof=1) / np.sqrt(n))],
            "per_seed": {"e": e.tolist(), "cv": cv.tolist(), "S": s.tolist()},
        })
    (Path(__file__).parent / "c3_reference.json").write_text(json.dumps(res, indent=1))
    for r in res["rungs"]:
        print(r["rung"], r["T"], r["e_per_spin"], r["cv_per_spin"], r["S_per_N"])


if __name__ == "__main__":
    main()
