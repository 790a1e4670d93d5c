"""Long runs of typed bodies with nuclei (the nuclei workload's physics at a
size the FP64 oracle runs in seconds), compared through per-type statistics
as the reference compares its GPU and CPU runs of this system
(docs/nucleus.md:71-117): the trajectories part after a while (README.md:245-248).

Workload: 8 replicas (seeds 1..8) of a 10 x 10 cell monolayer (packing 0.95,
modes 2..8) with one nucleus per cell (with_nuclei, ratio 0.4), 200 bodies;
cells: BodyType(v0 = 0.06, Dr = 0.01), OnTypes Tension(0.01) and
PerimeterElasticity(0.1, 3.7); nuclei: BodyType(Dr = 0.01, drag = parent),
OnTypes Bending(0.01); Overlap(J = 1, related = [0 -2; -2 0]) on all;
dt = 0.002, 10000 steps, sampled every 500 steps. The oracle's replicas are
sampled on the host (unwrapped_positions, shape_index by type,
src/observables.jl:6-29); the device replicas run as one ensemble observed
on the device. Per type: the replica-mean MSD at every sample within 4
standard errors (+2 % of the value), the mean shape index of the second half
within 4 standard errors (+1e-3), and the 10/50/90 % quantiles of the
second half's shape index within 0.5 % (device quantiles from a 2000-bin
histogram over [3.5, 5.5]).
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
NTYPES = 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def replica(seed):
    cells, box = F.hexagonal_monolayer(10, 10, packing=0.95, maxmode=8, seed=seed)
    b = F.with_nuclei(cells, ratio=0.4)
    types = (F.BodyType(v0=0.06, Dr=0.01, muab=0.1), F.BodyType(Dr=0.01, muab=0.1, drag="parent"))
    energies = (F.OnTypes(F.Tension(0.01), 1), F.OnTypes(F.PerimeterElasticity(0.1, 3.7), 1),
                F.OnTypes(F.Bending(0.01), 2),
                F.Overlap([[1.0, 1.0], [1.0, 1.0]], related=[[0.0, -2.0], [-2.0, 0.0]]))
    return F.Replica(b, box, types=types, energies=energies, dt=0.002, seed=seed, skin=0.3)


@pytest.fixture(scope="module")
def oracle_runs(oracle_engine):
    """Per replica: MSD [samples x types], and per type the shape indices of
    its bodies at every sample of the second half."""
    import oracle

    def one(seed):
        rp = replica(seed)
        s = F.Simulation(rp.bodies, rp.box, types=rp.types, energies=rp.energies, dt=rp.dt, seed=rp.seed,
                         skin=rp.skin, engine=oracle_engine)
        typ = rp.bodies.type
        r0 = F.unwrapped_positions(s)
        msd, p_half = [], [[] for _ in range(NTYPES)]
        for c in range(STEPS // EVERY):
            F.step(s, EVERY)
            d2 = np.sum((F.unwrapped_positions(s) - r0) ** 2, axis=0)
            msd.append([float(np.mean(d2[typ == t + 1])) for t in range(NTYPES)])
            if c >= STEPS // EVERY // 2:
                for t in range(NTYPES):
                    p_half[t].append(F.shape_index(s, type=t + 1))
        return np.array(msd), [np.concatenate(p) for p in p_half]

    say(f"oracle: 8 nucleated replicas x {STEPS} steps (FP64)")
    oracle.set_threads(1)
    with ThreadPoolExecutor(max_workers=8) as pool:
        return list(pool.map(one, SEEDS))


def device_runs(precision):
    ens = F.Ensemble([replica(s) for s in SEEDS], precision=precision)
    nck = STEPS // EVERY
    ens.observer_start(nbins=NBINS, prange=PRANGE, max_samples=nck)
    ens.run_observed(STEPS // 2, EVERY)
    hist_half = [o.hist.copy() for o in ens.observations()]
    ens.run_observed(STEPS - STEPS // 2, EVERY)
    out = []
    for o, h0 in zip(ens.observations(), hist_half):
        assert len(o.clock) == nck
        out.append((o.msd, o.p_mean[nck // 2:].mean(axis=0), (o.hist - h0)))
    ens.close()
    return out


def hist_quantiles(hist, qs):
    h = hist.astype(np.float64)
    total = h.sum()
    cum = h[0] + np.concatenate([[0.0], np.cumsum(h[1:-1])])
    at = np.asarray(qs) * total
    assert np.all(at > cum[0]) and np.all(at <= cum[-1])
    edges = PRANGE[0] + (PRANGE[1] - PRANGE[0]) * np.arange(NBINS + 1) / NBINS
    return np.interp(at, cum, edges)


@pytest.mark.parametrize("precision", ["mixed_fp16", "fp32"])
def test_nuclei_statistics_match_oracle(oracle_runs, precision):
    say("nuclei statistics", precision)
    dev = device_runs(precision)
    n = len(SEEDS)
    for t, name in enumerate(("cells", "nuclei")):
        m_o = np.array([r[0][:, t] for r in oracle_runs])
        m_d = np.array([r[0][:, t] for r in dev])
        for c in range(m_o.shape[1]):
            se = np.sqrt(m_o[:, c].var(ddof=1) / n + m_d[:, c].var(ddof=1) / n)
            if c % 5 == 4:
                say(f"  {name} MSD at step {(c + 1) * EVERY}: oracle {m_o[:, c].mean():.5f}, "
                    f"device {m_d[:, c].mean():.5f} (se {se:.5f})")
            assert abs(m_o[:, c].mean() - m_d[:, c].mean()) <= 4 * se + 0.02 * m_o[:, c].mean()
        p_o = np.array([r[1][t].mean() for r in oracle_runs])
        p_d = np.array([r[1][t] for r in dev])
        se = np.sqrt(p_o.var(ddof=1) / n + p_d.var(ddof=1) / n)
        say(f"  {name} <p>: oracle {p_o.mean():.5f}, device {p_d.mean():.5f} (se {se:.5f})")
        assert abs(p_o.mean() - p_d.mean()) <= 4 * se + 1e-3
        qo = np.quantile(np.concatenate([r[1][t] for r in oracle_runs]), [0.1, 0.5, 0.9])
        qd = hist_quantiles(np.sum([r[2][t] for r in dev], axis=0), [0.1, 0.5, 0.9])
        say(f"  {name} p quantiles 10/50/90 %: oracle {np.round(qo, 5)}, device {np.round(qd, 5)}")
        assert np.all(np.abs(qd - qo) <= 5e-3 * qo)
    m_o1, m_d1 = oracle_runs[0][0][-1], dev[0][0][-1]
    say(f"  final MSD difference of replica 1 (cells, nuclei): {np.abs(m_o1 - m_d1)}")
