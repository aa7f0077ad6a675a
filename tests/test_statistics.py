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
def _gpu(pyro_mod, prec, kind, n_or, R=256, n_burn=300, n_meas=2000, ssf_every=4, seed=31337, fused=False):
    ctx = pyro_mod.Context.pyrochlore(L, n_replicas=R, precision=prec, seed=seed)
    ctx.define_J1J2J3(*J)
    ctx.setup()
    ctx.set_fused_passes(fused)
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
    "prec,kind,n_or,fused",
    [(32, 0, 1, False), (32, 0, 2, False), (32, 1, 1, False), (32, 2, 1, False), (64, 0, 1, False),
     (64, 1, 1, False), (32, 0, 2, True), (32, 1, 1, True)],
    ids=["fp32-metropolis+or", "fp32-metropolis+2or", "fp32-lifted+or", "fp32-lifted_rot+or",
         "fp64-metropolis+or", "fp64-lifted+or", "fp32-metropolis+2or-fused", "fp32-lifted+or-fused"],
)
def test_sampled_observables_match_reference_sampler(gpu, prec, kind, n_or, fused):
    """The fused cases run the per-colour launches in the fused pass order
    (pyro_set_fused_passes), the order of the tempering workload."""
    ref = _reference(kind, n_or)
    got = _gpu(gpu, prec, kind, n_or, fused=fused)
    print(f"GPU FP{prec} kind {kind} n_or {n_or}{' fused order' if fused else ''} vs reference sampler:")
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


# ---------------------------------------------------------------------------
# Tiled sweeps with a site's passes back to back (pyro_set_fused_passes):
# a different sequential order of the same single-site updates, so the same
# stationary distribution.
# ---------------------------------------------------------------------------
def _tiled_ensemble(pyro_mod, strategy, fused, L8=8, R=128, T8=0.4, kind=1, n_or=1, n_burn=300, n_meas=1000,
                    seed=4242):
    ctx = pyro_mod.Context.pyrochlore(L8, n_replicas=R, precision=pyro_mod.FP32, seed=seed)
    ctx.define_J1J2J3(*J)
    ctx.setup()
    ctx.set_strategy(strategy)
    ctx.set_fused_passes(fused)
    ctx.set_temperatures(np.full(R, T8))
    ctx.randomize(seed + 1)
    ctx.ssf_configure(("xx", "yy", "zz"), n_slots=R, store_error_term=False)
    ctx.sweep(kind, n_or, n_burn, count=False)
    ctx.anneal_schedule(np.full(n_meas, T8))
    ctx.anneal_steps(kind, n_or, 1, True, n_meas)
    e = ctx.anneal_history() * L8 ** 3  # [R][n_meas], total energy
    N8 = 16 * L8 ** 3
    s_runs = np.zeros((R, len(KS)))
    for r in range(R):
        S, _, n = ctx.ssf_read(r, with_sq=False)
        assert n == n_meas
        s_runs[r] = S[:, list(KS), 0].sum(0) / n / N8
    ctx.close()
    E, E2 = e.mean(1), (e * e).mean(1)
    R_ = len(E)
    cv = lambda a, b: (b - a * a) / (N8 * T8 * T8)  # noqa: E731
    loo = np.array([cv((E.sum() - E[r]) / (R_ - 1), (E2.sum() - E2[r]) / (R_ - 1)) for r in range(R_)])
    out = {"e": ((E / N8).mean(), (E / N8).std(ddof=1) / np.sqrt(R_)),
           "cv": (cv(E.mean(), E2.mean()), np.sqrt((R_ - 1) / R_ * np.sum((loo - loo.mean()) ** 2)))}
    for j, K in enumerate(KS):
        out[f"S{K}"] = (s_runs[:, j].mean(), s_runs[:, j].std(ddof=1) / np.sqrt(R_))
    return out


def test_fused_tiled_order_samples_the_same_ensemble(gpu):
    """L = 8, lifted + OR at T = 0.4: the tiled sweep with fused passes, the
    tiled sweep pass by pass and per-colour launches (the order the
    reference-sampler tests cover) give the same <E>/N, C_v/N and S(q)
    within 3 sigma of the replica-to-replica errors."""
    fused = _tiled_ensemble(gpu, 5, True)
    tiled = _tiled_ensemble(gpu, 5, False)
    colour = _tiled_ensemble(gpu, 3, False)
    print("fused tiled vs tiled pass by pass:")
    bad = _compare(fused, tiled)
    print("fused tiled vs per-colour launches:")
    bad += _compare(fused, colour)
    assert not bad, f"outside {N_SIGMA} sigma: {bad}"
    assert fused["e"][1] < 1e-3


def test_fused_tiled_order_langevin_law(gpu):
    """Spins decoupled up to J1 = 1e-9 in a field B along z: every
    single-site update samples exp(S.B / T) on the sphere, so
    <S_z> = coth(|B|/T) - T/|B| exactly (the Langevin function), whatever
    the order; lifted + OR with fused passes on 4x4x4 tiles."""
    L8, R, Tl, Bz = 8, 64, 0.5, 1.0
    ctx = gpu.Context.pyrochlore(L8, n_replicas=R, precision=gpu.FP32, seed=99)
    ctx.define_J1J2J3(1e-9)
    ctx.setup()
    ctx.set_field(np.array([0.0, 0.0, Bz]))
    ctx.set_strategy(5)
    ctx.set_fused_passes(True)
    ctx.set_temperatures(np.full(R, Tl))
    ctx.randomize(3)
    ctx.sweep(gpu.LIFTED, 1, 200, count=False)
    n_meas = 500
    ctx.anneal_schedule(np.full(n_meas, Tl))
    ctx.anneal_steps(gpu.LIFTED, 1, 1, False, n_meas)
    mz = -ctx.anneal_history().mean(1) / 16.0 / Bz  # per replica <S_z> (E/cell = -B . sum S / L^3)
    x = Bz / Tl
    want = 1.0 / np.tanh(x) - 1.0 / x
    err = mz.std(ddof=1) / np.sqrt(R)
    print(f"<S_z> = {mz.mean():.6f} +- {err:.6f}, Langevin {want:.6f}")
    assert abs(mz.mean() - want) < 3 * err + 1e-6
    assert err < 1e-3
