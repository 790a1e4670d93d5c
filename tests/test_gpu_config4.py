"""The batched-ensemble workload at its own size: 64 replicas with N
log-uniform in [1024, 16384] (367774 bodies), packing U[0.9, 1], v0 swept
over 0.02..0.2, Dr over {0.005, 0.01, 0.02}, Tension(η ~ U[0.005, 0.05]) +
Overlap(1.0), modes 2..6, FP32, Philox key = replica id (the replica set of
bench.py's ensemble leg), as the reference's concurrent simulations
(README.md:249-261, test/test_cuda.jl:95-104): launch shape from the
autotuner, 200 steps as one ensemble; the smallest and the largest replica,
the fastest (v0 = 0.2) and one with Dr = 0.02 end bitwise where they end when
run alone (with the lane count per body the autotuner chose), and the smallest replica's end state has gradients within the
force tolerance of the FP64 oracle's on the same state."""
import sys

import numpy as np
import pytest

import fcc_hip as F
from helpers import force_scales, plain_force_errors, rel_err

pytestmark = pytest.mark.gpu

STEPS = 200


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def state(s):
    b, c = s.bodies, s.counters()
    return (b.x, b.y, b.theta, b.coef, b.imgx, b.imgy), (c.step, c.t, c.clock, c.builds)


def test_config4_workload(oracle_engine):
    import bench
    import oracle
    reps = bench.ensemble_replicas()
    sizes = np.array([len(rp.bodies) for rp in reps])
    assert len(reps) == 64 and sizes.min() >= 1024 and sizes.max() <= 16384 + 256
    v0 = np.array([rp.types[0].v0 for rp in reps])
    Dr = np.array([rp.types[0].Dr for rp in reps])
    picks = [int(np.argmin(sizes)), int(np.argmax(sizes)), int(np.argmax(v0))]
    picks.append(next(r for r in range(len(reps)) if Dr[r] == 0.02 and r not in picks))
    say(f"config 4: {len(reps)} replicas, {sizes.sum()} bodies; checked alone: {picks} "
        f"(N {sizes[picks].tolist()}, v0 {np.round(v0[picks], 3).tolist()}, Dr {Dr[picks].tolist()})")
    ens = F.Ensemble(reps, precision="fp32")
    launch, ms = ens.autotune(10)
    say(f"  autotuned launch {launch}, {ms:.4f} ms per ensemble step")
    ens.run(STEPS)
    builds = [ens.members[r].counters().builds for r in picks]
    say(f"  ensemble ran {STEPS} steps; list builds of the checked replicas {builds}")
    for r in picks:
        rp = reps[r]
        alone = F.Simulation(rp.bodies, rp.box, types=rp.types, energies=rp.energies, dt=rp.dt, seed=rp.seed,
                             skin=rp.skin, precision="fp32", two_lanes=launch["lanes"] == 2)
        F.run(alone, STEPS)
        (sa, ca), (se, ce) = state(alone), state(ens.members[r])
        for u, w in zip(sa, se):
            assert np.array_equal(u, w), f"replica {r}"
        assert ca == ce, f"replica {r}"
        alone.close()
    assert max(builds) > 0

    # the FP64 oracle on the smallest replica's end state
    r = picks[0]
    rp, end = reps[r], ens.members[r].bodies
    ens.close()
    kw = dict(types=rp.types, energies=rp.energies, dt=rp.dt, seed=rp.seed, skin=rp.skin, debug=True)
    oracle.set_threads(16)
    ref_sim = F.refresh(F.Simulation(end, rp.box, engine=oracle_engine, **kw))
    dev_sim = F.refresh(F.Simulation(end, rp.box, precision="fp32", **kw))
    ref, dev = ref_sim.grad, dev_sim.grad
    flips = dev.P != ref.P
    nflip, ncontact = int(np.count_nonzero(flips)), int(np.count_nonzero(ref.P))
    ok = ~np.any(flips, axis=0)
    fs, ts = force_scales(ref_sim, ref)
    ef = np.maximum(rel_err(dev.gx, ref.gx, fs), rel_err(dev.gy, ref.gy, fs))[ok]
    et = rel_err(dev.gth, ref.gth, ts)[ok]
    ec = np.max(np.abs(dev.gcoef[1:, ok] - ref.gcoef[1:, ok])) / np.max(np.abs(ref.gcoef[1:, ok]))
    med, q99, pmax, field = plain_force_errors(dev, ref, ok)
    say(f"  replica {r} (N {sizes[r]}) vs FP64 oracle: flips {nflip} of {ncontact} contact nodes; force "
        f"{ef.max():.2e}, torque {et.max():.2e}, gcoef {ec:.2e}; plain |dF|/|F| median {med:.2e}, max "
        f"{pmax:.2e}; ||dF||/||F|| {field:.2e}")
    assert ncontact > 1000
    assert nflip <= max(2, ncontact // 1000)
    assert ok.sum() >= 0.99 * ok.size
    assert ef.max() <= 1e-4 and et.max() <= 1e-4 and ec <= 1e-4 and field <= 1e-4
