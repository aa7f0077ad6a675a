"""The anneal configurations of the workload list, run at their full sizes
through the batch driver (one context per configuration, one replica per
seed), with checks that do not depend on this implementation:

- plain anneal, FP64, L = 4, J1 = 1 (T 10 -> 0.05 over 200 log steps,
  n_burn_in 16, n_sweep 16, n_sample 128, Metropolis + OR): the classical
  Heisenberg pyrochlore antiferromagnet has E/spin -> -J1 with C/N -> 3/4
  as T -> 0 (Moessner and Chalker 1998), so at T_cold = 0.05 the mean
  E/spin over seeds is -1 + 0.75 * 0.05 up to O(T^2);
- FP64 and FP32 side by side, L = 16, J1 = 1, J2 = 0.1 (T 10 -> 0.02 over
  100 log steps, n_sweep 16): the seed-averaged energy of the two
  precisions agrees within 3 sigma at every step of the anneal.
"""
import numpy as np
import pytest


def _batch(gpu, L, J, n_seeds, precision, **kw):
    from pyro_cmc_amd import batch

    items = [batch.AnnealItem(L, J) for _ in range(n_seeds)]
    res, _ = batch.run_anneal_batch(items, precision=precision, **kw)
    return res


@pytest.mark.gpu
def test_plain_anneal_fp64_low_temperature_energy(gpu):
    res = _batch(gpu, 4, (1.0, 0.0, 0.0, 1.0), 32, gpu.FP64, T_hot=10.0, T_cold=0.05, n_steps=200,
                 n_sweep=16, n_burn_in=16, n_sample=128, seed=1)
    e = np.array([r.E_final for r in res]) / 16.0  # per spin
    want = -1.0 + 0.75 * 0.05
    sem = e.std(ddof=1) / np.sqrt(len(e))
    # statistical error plus the O(T^2) remainder of the low-T expansion
    assert abs(e.mean() - want) < 3 * sem + 0.05 ** 2, (e.mean(), sem)


@pytest.mark.gpu
def test_fp32_and_fp64_anneals_agree(gpu):
    kw = dict(T_hot=10.0, T_cold=0.02, n_steps=100, n_sweep=16, n_burn_in=16, n_sample=64, seed=2)
    r64 = _batch(gpu, 16, (1.0, 0.1, 0.0, 1.0), 16, gpu.FP64, **kw)
    r32 = _batch(gpu, 16, (1.0, 0.1, 0.0, 1.0), 16, gpu.FP32, **kw)
    e64 = np.array([r.E for r in r64])  # [seed][step], per cell
    e32 = np.array([r.E for r in r32])
    se = np.sqrt(e64.var(0, ddof=1) / len(e64) + e32.var(0, ddof=1) / len(e32))
    z = np.abs(e64.mean(0) - e32.mean(0)) / se
    # 100 steps: allow the expected few 2-sigma steps, none beyond 4 sigma
    assert np.mean(z > 3) < 0.05 and z.max() < 4.5, (z.max(), np.argmax(z))
    f64 = np.array([r.E_final for r in r64]) / 16
    f32 = np.array([r.E_final for r in r32]) / 16
    se_f = np.sqrt(f64.var(ddof=1) / len(f64) + f32.var(ddof=1) / len(f32))
    assert abs(f64.mean() - f32.mean()) < 3 * se_f
