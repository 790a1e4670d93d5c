"""Long runs: the trajectories of the FP64 oracle and of the device paths part
after a few hundred steps (the contact dynamics amplifies round-off,
README.md:245-248), so they are compared through statistics of an ensemble of
replicas: the shape-index distribution (src/observables.jl:6-18) and the mean
squared displacement of the unwrapped positions (src/observables.jl:24-29).
The oracle's replicas are sampled on the host with those functions; the device
replicas run as one ensemble observed on the device (fcc_ensemble_run_observed,
whose samples equal the host functions on the same state:
test_gpu_observables.py).

Workload: 8 replicas (seeds 1..8) of a 12 x 12 monolayer (packing 0.95, modes
2..6), BodyType(v0 = 0.3, Dr = 0.1), Tension(0.01) + Overlap(1.0), dt = 0.005,
10000 steps, observed every 500 steps. Tolerances: replica means within 4
standard errors (plus 2 % of the value for the MSD, 1e-3 for the shape index),
and the 10/50/90 % quantiles of the pooled shape index of the second half
within 0.5 % (device quantiles from a 2000-bin histogram over [3.5, 5.5],
bin width 1e-3).
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import fcc_hip as F

pytestmark = pytest.mark.gpu

SEEDS = range(1, 9)
STEPS, EVERY = 10000, 500
NBINS, PRANGE = 2000, (3.5, 5.5)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def replica(seed):
    b, box = F.hexagonal_monolayer(12, 12, packing=0.95, maxmode=6, seed=seed)
    return F.Replica(b, box, types=(F.BodyType(v0=0.3, Dr=0.1),),
                     energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005, seed=seed, skin=0.3)


@pytest.fixture(scope="module")
def oracle_runs(oracle_engine):
    """Per replica: the MSD at every sample and the shape indices of every
    body at every sample of the second half."""
    import oracle

    def one(seed):
        rp = replica(seed)
        s = F.Simulation(rp.bodies, rp.box, types=rp.types, energies=rp.energies, dt=rp.dt, seed=rp.seed,
                         skin=rp.skin, engine=oracle_engine)
        r0 = F.unwrapped_positions(s)
        msd, p_half = [], []
        for c in range(STEPS // EVERY):
            F.step(s, EVERY)
            r = F.unwrapped_positions(s)
            msd.append(float(np.mean(np.sum((r - r0) ** 2, axis=0))))
            if c >= STEPS // EVERY // 2:
                p_half.append(F.shape_index(s))
        return np.array(msd), np.concatenate(p_half)

    say(f"oracle: 8 replicas x {STEPS} steps (FP64)")
    oracle.set_threads(1)  # one host thread per replica
    with ThreadPoolExecutor(max_workers=8) as pool:
        return list(pool.map(one, SEEDS))


def device_runs(precision):
    """The same per replica from the device observer: MSD series, mean shape
    index over the second half's samples, and the histogram of the second
    half (the cumulative histogram at the end minus the one at the half)."""
    ens = F.Ensemble([replica(s) for s in SEEDS], precision=precision)
    nck = STEPS // EVERY
    ens.observer_start(nbins=NBINS, prange=PRANGE, max_samples=nck)
    ens.run_observed(STEPS // 2, EVERY)
    hist_half = [o.hist.copy() for o in ens.observations()]
    ens.run_observed(STEPS - STEPS // 2, EVERY)
    out = []
    for o, h0 in zip(ens.observations(), hist_half):
        assert len(o.clock) == nck
        out.append((o.msd[:, 0], o.p_mean[nck // 2:, 0].mean(), (o.hist - h0)[0]))
    return out


def hist_quantiles(hist, qs):
    """Quantiles of a histogram (bin 0: below PRANGE, last bin: at or above),
    linear within a bin; every quantile asked for must fall inside PRANGE."""
    h = hist.astype(np.float64)
    total = h.sum()
    cum = h[0] + np.concatenate([[0.0], np.cumsum(h[1:-1])])
    at = np.asarray(qs) * total
    assert np.all(at > cum[0]) and np.all(at <= cum[-1])
    edges = PRANGE[0] + (PRANGE[1] - PRANGE[0]) * np.arange(NBINS + 1) / NBINS
    return np.interp(at, cum, edges)


@pytest.mark.parametrize("precision", ["fp32", "mixed_fp16", "mixed_bf16"])
def test_long_run_statistics_match_oracle(oracle_runs, precision):
    say("statistics", precision)
    dev = device_runs(precision)
    n = len(SEEDS)
    m_o = np.array([r[0] for r in oracle_runs])  # replicas x samples
    m_d = np.array([r[0] for r in dev])
    for c in range(m_o.shape[1]):
        se = np.sqrt(m_o[:, c].var(ddof=1) / n + m_d[:, c].var(ddof=1) / n)
        if c % 4 == 3:
            say(f"  MSD at step {(c + 1) * EVERY}: oracle {m_o[:, c].mean():.4f}, device {m_d[:, c].mean():.4f} "
                f"(se {se:.4f})")
        assert abs(m_o[:, c].mean() - m_d[:, c].mean()) <= 4 * se + 0.02 * m_o[:, c].mean()
    p_o = np.array([r[1].mean() for r in oracle_runs])
    p_d = np.array([r[1] for r in dev])
    se = np.sqrt(p_o.var(ddof=1) / n + p_d.var(ddof=1) / n)
    say(f"  <p>: oracle {p_o.mean():.5f}, device {p_d.mean():.5f} (se {se:.5f})")
    assert abs(p_o.mean() - p_d.mean()) <= 4 * se + 1e-3
    pool_o = np.concatenate([r[1] for r in oracle_runs])
    qo = np.quantile(pool_o, [0.1, 0.5, 0.9])
    qd = hist_quantiles(np.sum([r[2] for r in dev], axis=0), [0.1, 0.5, 0.9])
    say(f"  p quantiles 10/50/90 %: oracle {np.round(qo, 5)}, device {np.round(qd, 5)}")
    assert np.all(np.abs(qd - qo) <= 5e-3 * qo)
    # the trajectories themselves have parted (otherwise this test would be a
    # trajectory comparison)
    say(f"  final MSD difference of replica 1: {abs(m_o[0, -1] - m_d[0, -1]):.3e}")
