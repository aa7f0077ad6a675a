"""Sampled observables of the GPU path against the reference sampler.

The GPU's counter RNG and colour-ordered sweeps cannot reproduce the
reference's Xoroshiro stream in storage order, so sampled observables are
compared statistically: <E>/N, C_v/N and S(q) must agree within 3 sigma.
  * reference side: the oracle in reference mode, which follows the
    reference's site order and draws (MC.cpp:111-303 with Xoroshiro128++ and
    the libstdc++ distributions), over independent seeds;
  * GPU side: independent replicas of one context at the same temperature.
Errors come from the spread over independent runs (replicas or seeds); C_v
uses a jackknife over them.  Model: pyrochlore L = 2 with the tempering
couplings J = (1, -0.05, 0.05), B = 0, T = 0.4 (cooperative paramagnet,
C_v well away from its limits).

The last test compares the FP32 ensemble with the FP64 one per observable
(the mixed-precision error statement).
"""
import functools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = 2
J = (1.0, -0.05, 0.05)
T = 0.4
N = 16 * L ** 3
KS = (0, 1, 7)  # K = (0,0,0), (0,0,1), (1,1,1)
N_SIGMA = 3.0


def _jackknife_cv(E, E2):
    """C_v / N from per-run means of E_tot and E_tot^2 (equal sample counts)."""
    R = len(E)
    cv = lambda e, e2: (e2 - e * e) / (N * T * T)  # noqa: E731
    full = cv(E.mean(), E2.mean())
    loo = np.array([cv((E.sum() - E[r]) / (R - 1), (E2.sum() - E2[r]) / (R - 1)) for r in range(R)])
    return full, np.sqrt((R - 1) / R * np.sum((loo - loo.mean()) ** 2))


def _summary(e_runs, E_runs, E2_runs, s_runs):
    """(value, sigma) per observable from independent runs."""
    R = len(e_runs)
    out = {"e": (e_runs.mean(), e_runs.std(ddof=1) / np.sqrt(R)), "cv": _jackknife_cv(E_runs, E2_runs)}
    for j, K in enumerate(KS):
        out[f"S{K}"] = (s_runs[:, j].mean(), s_runs[:, j].std(ddof=1) / np.sqrt(R))
    return out


@functools.lru_cache(maxsize=None)
def _reference(kind, n_or, n_seeds=32, n_burn=300, n_meas=2000, ssf_every=4):
    import pyoracle

    e_runs, E_runs, E2_runs, s_runs = [], [], [], []
    rng = np.random.default_rng(1000 + 10 * kind + n_or)
    for seed in range(n_seeds):
        o = pyoracle.Oracle.pyrochlore(L, *J, seed=7919 * seed + 17)
        v = rng.normal(size=(o.N, 3))
        o.set_spins(v / np.linalg.norm(v, axis=1, keepdims=True))
        for _ in range(n_burn):
            o.sweep_ref(kind, T, n_or)
        es, ss = [], []
        for t in range(n_meas):
            o.sweep_ref(kind, T, n_or)
            es.append(o.energy_per_cell() * L ** 3)
            if t % ssf_every == ssf_every - 1:
                s = o.ssf_sample(("xx", "yy", "zz"))
                ss.append(s[:, list(KS), 0].sum(0) / N)
        es = np.array(es)
        e_runs.append(es.mean() / N)
        E_runs.append(es.mean())
        E2_runs.append((es ** 2).mean())
        s_runs.append(np.mean(ss, 0))
    return _summary(np.array(e_runs), np.array(E_runs), np.array(E2_runs), np.array(s_runs))


@functools.lru_cache(maxsize=None)
def _gpu(pyro_mod, prec, kind, n_or, R=256, n_burn=300, n_meas=2000, ssf_every=4, seed=31337):
    ctx = pyro_mod.Context.pyrochlore(L, n_replicas=R, precision=prec, seed=seed)
    ctx.define_J1J2J3(*J)
    ctx.setup()
    ctx.set_temperatures(np.full(R, T))
    ctx.randomize(seed + 1)
    ctx.ssf_configure(("xx", "yy", "zz"), n_slots=R, store_error_term=False)
    ctx.sweep(kind, n_or, n_burn)
    E = np.zeros(R)
    E2 = np.zeros(R)
    for t in range(n_meas):
        ctx.sweep(kind, n_or, 1)
        e = ctx.energy_per_cell() * L ** 3
        E += e
        E2 += e * e
        if t % ssf_every == ssf_every - 1:
            ctx.ssf_sample()
    E /= n_meas
    E2 /= n_meas
    s_runs = np.zeros((R, len(KS)))
    for r in range(R):
        S, _, n = ctx.ssf_read(r, with_sq=False)
        assert n == n_meas // ssf_every
        s_runs[r] = S[:, list(KS), 0].sum(0) / n / N
    return _summary(E / N, E, E2, s_runs)


def _compare(a, b, n_sigma=N_SIGMA):
    bad = []
    for k in a:
        (va, sa), (vb, sb) = a[k], b[k]
        s = np.hypot(sa, sb)
        print(f"  {k:>4}: {va:+.6f} +- {sa:.6f}  vs  {vb:+.6f} +- {sb:.6f}   ({(va - vb) / s:+.2f} sigma)")
        if not abs(va - vb) <= n_sigma * s:
            bad.append(k)
    return bad


@pytest.mark.parametrize(
    "prec,kind,n_or",
    [(32, 0, 1), (32, 0, 2), (32, 1, 1), (32, 2, 1), (64, 0, 1), (64, 1, 1)],
    ids=["fp32-metropolis+or", "fp32-metropolis+2or", "fp32-lifted+or", "fp32-lifted_rot+or",
         "fp64-metropolis+or", "fp64-lifted+or"],
)
def test_sampled_observables_match_reference_sampler(gpu, prec, kind, n_or):
    ref = _reference(kind, n_or)
    got = _gpu(gpu, prec, kind, n_or)
    print(f"GPU FP{prec} kind {kind} n_or {n_or} vs reference sampler:")
    bad = _compare(got, ref)
    assert not bad, f"outside {N_SIGMA} sigma: {bad}"
    # the error bars must be small enough for the comparison to mean something
    assert got["e"][1] < 2e-3 and ref["e"][1] < 5e-3
    assert ref["cv"][1] < 0.1 * abs(ref["cv"][0])


@pytest.mark.parametrize("kind,n_or", [(0, 1), (1, 1)])
def test_fp32_against_fp64_ensemble(gpu, kind, n_or):
    """Per-observable FP32 error against the all-FP64 GPU path (same seeds,
    same sweeps): differences stay within the statistical error."""
    a = _gpu(gpu, 32, kind, n_or)
    b = _gpu(gpu, 64, kind, n_or)
    print(f"FP32 vs FP64, kind {kind} n_or {n_or}:")
    bad = _compare(a, b)
    assert not bad, f"FP32 differs from FP64 beyond {N_SIGMA} sigma: {bad}"
