"""Long runs: the trajectories of the FP64 oracle and of the device paths part
after a few hundred steps (the contact dynamics amplifies round-off,
README.md:245-248), so they are compared through statistics of an ensemble of
replicas: the shape-index distribution (src/observables.jl:6-18) and the mean
squared displacement of the unwrapped positions (src/observables.jl:24-29).

Workload: 8 replicas (seeds 1..8) of a 12 x 12 monolayer (packing 0.95, modes
2..6), BodyType(v0 = 0.3, Dr = 0.1), Tension(0.01) + Overlap(1.0), dt = 0.005,
3000 steps, observed every 500 steps. The device replicas run as one ensemble.
Tolerances: replica means within 4 standard errors (plus 2 % of the value for
the MSD, 1e-3 for the shape index), and the 10/50/90 % quantiles of the pooled
shape index of the second half within 0.5 %.
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import fcc_hip as F

pytestmark = pytest.mark.gpu

SEEDS = range(1, 9)
STEPS, EVERY = 3000, 500


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def replica(seed):
    b, box = F.hexagonal_monolayer(12, 12, packing=0.95, maxmode=6, seed=seed)
    return F.Replica(b, box, types=(F.BodyType(v0=0.3, Dr=0.1),),
                     energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005, seed=seed, skin=0.3)


def observe(sim, r0):
    r = F.unwrapped_positions(sim)
    return F.shape_index(sim), float(np.mean(np.sum((r - r0) ** 2, axis=0)))


@pytest.fixture(scope="module")
def oracle_runs(oracle_engine):
    import oracle

    def one(seed):
        rp = replica(seed)
        s = F.Simulation(rp.bodies, rp.box, types=rp.types, energies=rp.energies, dt=rp.dt, seed=rp.seed,
                         skin=rp.skin, engine=oracle_engine)
        r0 = F.unwrapped_positions(s)
        out = []
        for _ in range(STEPS // EVERY):
            F.step(s, EVERY)
            out.append(observe(s, r0))
        return out

    say("oracle: 8 replicas x 3000 steps (FP64)")
    oracle.set_threads(1)  # one host thread per replica
    with ThreadPoolExecutor(max_workers=8) as pool:
        return list(pool.map(one, SEEDS))


def device_runs(precision):
    ens = F.Ensemble([replica(s) for s in SEEDS], precision=precision)
    r0 = [F.unwrapped_positions(m) for m in ens.members]
    out = [[] for _ in SEEDS]
    for _ in range(STEPS // EVERY):
        ens.run(EVERY)
        for q, m in enumerate(ens.members):
            out[q].append(observe(m, r0[q]))
    return out


def summary(runs):
    """Per checkpoint: replica MSDs; second half: replica mean shape index and
    the pooled shape-index values."""
    nck = len(runs[0])
    msd = np.array([[runs[r][c][1] for r in range(len(runs))] for c in range(nck)])
    half = range(nck // 2, nck)
    pmean = np.array([np.mean([runs[r][c][0].mean() for c in half]) for r in range(len(runs))])
    pooled = np.concatenate([runs[r][c][0] for r in range(len(runs)) for c in half])
    return msd, pmean, pooled


@pytest.mark.parametrize("precision", ["fp32", "mixed_fp16", "mixed_bf16"])
def test_long_run_statistics_match_oracle(oracle_runs, precision):
    say("statistics", precision)
    dev = device_runs(precision)
    m_o, p_o, pool_o = summary(oracle_runs)
    m_d, p_d, pool_d = summary(dev)
    n = len(SEEDS)
    for c in range(m_o.shape[0]):
        se = np.sqrt(m_o[c].var(ddof=1) / n + m_d[c].var(ddof=1) / n)
        say(f"  MSD at step {(c + 1) * EVERY}: oracle {m_o[c].mean():.4f}, device {m_d[c].mean():.4f} "
            f"(se {se:.4f})")
        assert abs(m_o[c].mean() - m_d[c].mean()) <= 4 * se + 0.02 * m_o[c].mean()
    se = np.sqrt(p_o.var(ddof=1) / n + p_d.var(ddof=1) / n)
    say(f"  <p>: oracle {p_o.mean():.5f}, device {p_d.mean():.5f} (se {se:.5f})")
    assert abs(p_o.mean() - p_d.mean()) <= 4 * se + 1e-3
    qo, qd = np.quantile(pool_o, [0.1, 0.5, 0.9]), np.quantile(pool_d, [0.1, 0.5, 0.9])
    say(f"  p quantiles 10/50/90 %: oracle {np.round(qo, 5)}, device {np.round(qd, 5)}")
    assert np.all(np.abs(qd - qo) <= 5e-3 * qo)
    # the trajectories themselves have parted (otherwise this test would be a
    # trajectory comparison)
    say(f"  final MSD difference of replica 1: {abs(m_o[-1][0] - m_d[-1][0]):.3e}")
