"""Checkpoint and resume (fcc_checkpoint_save / fcc_checkpoint_load): a run
interrupted, saved, loaded into a new simulation of the same configuration
and continued must end bitwise where the uninterrupted run ends — positions,
orientations, shapes, image counters, device clock (Philox draw index),
step, time and list builds. The uninterrupted run is itself pinned to the
FP64 oracle by test_gpu_parity.py; a resume that lost the clock or the Verlet
reference would draw other noise or rebuild at other steps and part from it.
"""
import sys

import numpy as np
import pytest

import fcc_hip as F

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(precision, seed=3, n=10, maxmode=6, skin=0.05):
    b, box = F.hexagonal_monolayer(n, n, packing=0.95, maxmode=maxmode, seed=seed)
    return (b, box), dict(types=(F.BodyType(v0=2.0, Dr=0.5),), energies=(F.Tension(0.01), F.Overlap(1.0)),
                          precision=precision, seed=seed, skin=skin)


def state(s):
    b, c = s.bodies, s.counters()
    return (b.x, b.y, b.theta, b.coef, b.imgx, b.imgy), (c.step, c.t, c.clock, c.builds)


def assert_same(a, b):
    (sa, ca), (sb, cb) = state(a), state(b)
    for u, v in zip(sa, sb):
        assert np.array_equal(u, v)
    assert ca == cb


@pytest.mark.parametrize("precision,maxmode", [("fp64", 6), ("fp32", 6), ("mixed_fp16", 8), ("mixed_bf16", 6)])
def test_resume_continues_bitwise(precision, maxmode):
    say("resume", precision, maxmode)
    (b, box), kw = make(precision, maxmode=maxmode)
    full = F.Simulation(b, box, **kw)
    F.run(full, 150)
    F.step(full, 7)
    F.run(full, 143)
    part = F.Simulation(b, box, **kw)
    F.run(part, 150)
    F.step(part, 7)
    blob = part.checkpoint()
    part.close()
    resumed = F.Simulation(b, box, **kw)
    resumed.restore(blob)
    assert resumed.counters().clock == 157
    F.run(resumed, 143)
    assert_same(full, resumed)
    assert full.counters().builds > 10


def test_resume_without_the_checkpoint_differs():
    # the control: bodies alone (the reference's only way to resume) restart
    # the clock, so the continued run draws other noise
    (b, box), kw = make("fp32")
    full = F.Simulation(b, box, **kw)
    F.run(full, 100)
    mid = F.Simulation(b, box, **kw)
    F.run(mid, 50)
    fresh = F.Simulation(mid.bodies, box, **kw)
    F.run(fresh, 50)
    assert not np.array_equal(full.bodies.x, fresh.bodies.x)


def test_member_checkpoint_resumes_alone_and_in_an_ensemble():
    say("ensemble member checkpoints")
    reps = []
    for s in (1, 2):
        (b, box), kw = make("fp32", seed=s, n=8 + 4 * s)
        reps.append(F.Replica(b, box, types=kw["types"], energies=kw["energies"], seed=s, skin=kw["skin"]))
    ens = F.Ensemble(reps, precision="fp32")
    ens.run(60)
    blobs = [m.checkpoint() for m in ens.members]
    ens.run(40)
    for rp, blob, m in zip(reps, blobs, ens.members):
        alone = F.Simulation(rp.bodies, rp.box, types=rp.types, energies=rp.energies, dt=rp.dt, seed=rp.seed,
                             skin=rp.skin, precision="fp32")
        alone.restore(blob)
        F.run(alone, 40)
        assert_same(alone, m)
    # and back into the ensemble's replicas: the same 40 steps again
    ref = [state(m) for m in ens.members]
    for m, blob in zip(ens.members, blobs):
        m.restore(blob)
    ens.run(40)
    for m, (sr, cr) in zip(ens.members, ref):
        (sm, cm) = state(m)
        for u, v in zip(sm, sr):
            assert np.array_equal(u, v)
        assert cm == cr


def test_checkpoint_checks():
    (b, box), kw = make("fp32")
    s = F.Simulation(b, box, **kw)
    F.run(s, 20)
    blob = s.checkpoint()
    with pytest.raises(F.DimensionMismatch):
        F.Simulation(b, box, **{**kw, "precision": "fp64"}).restore(blob)
    with pytest.raises(F.ArgumentError):
        F.Simulation(b, box, **{**kw, "seed": 99}).restore(blob)
    with pytest.raises(F.DimensionMismatch):
        F.Simulation(b, box, **kw).restore(blob[: len(blob) // 2])
    with pytest.raises(F.ArgumentError):
        F.Simulation(b, box, **kw).restore(b"X" + blob[1:])
    two = {**kw, "types": (F.BodyType(v0=2.0, Dr=0.5), F.BodyType(v0=2.0, Dr=0.5))}
    with pytest.raises(F.ArgumentError):
        F.Simulation(b, box, **two).restore(blob)  # another number of types
    b2 = b.copy()
    b2.type[::2] = 2
    s2 = F.Simulation(b, box, **two)
    blob2 = s2.checkpoint()
    with pytest.raises(F.ArgumentError):
        F.Simulation(b2, box, **two).restore(blob2)  # other body types


def test_resumed_run_matches_oracle_fp64(oracle_engine):
    # The resumed run against the FP64 oracle's uninterrupted one (device
    # clock, src/simulation.jl:99; Verlet reference, src/neighbors.jl:234-240):
    # 150 steps, checkpoint, a fresh simulation restored from it, 50 more
    # steps; the oracle runs 200 steps. Rotational and translational noise
    # on, thin skin: the continued run draws the noise of draws 151..200 and
    # rebuilds where the uninterrupted run does.
    say("resume vs oracle")
    b, box = F.hexagonal_monolayer(8, 8, packing=0.95, maxmode=6, seed=5)
    kw = dict(types=(F.BodyType(v0=0.3, Dr=0.5, Dt=0.01),), energies=(F.Tension(0.01), F.Overlap(1.0)),
              seed=5, skin=0.05)
    ref = F.Simulation(b, box, engine=oracle_engine, **kw)
    F.step(ref, 200)
    part = F.Simulation(b, box, precision="fp64", **kw)
    F.run(part, 150)
    blob = part.checkpoint()
    part.close()
    dev = F.Simulation(b, box, precision="fp64", **kw)
    dev.restore(blob)
    F.run(dev, 50)
    bd, br = dev.bodies, ref.bodies
    cd, cr = dev.counters(), ref.counters()
    say(f"  builds device {cd.builds}, oracle {cr.builds}; max |dx| {np.max(np.abs(bd.x - br.x)):.2e}")
    np.testing.assert_allclose(bd.x, br.x, rtol=1e-8)
    np.testing.assert_allclose(bd.y, br.y, rtol=1e-8)
    np.testing.assert_allclose(bd.theta, br.theta, rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(bd.coef, br.coef, rtol=1e-8, atol=1e-12)
    assert np.array_equal(bd.imgx, br.imgx) and np.array_equal(bd.imgy, br.imgy)
    assert cd.clock == cr.clock == 200 and cd.step == cr.step == 200
    assert cd.builds == cr.builds and cr.builds > 5
