"""Device-side observations (fcc_observer_*, fcc_observe, fcc_run_observed):
each sample must equal what the reference's host observables give on the
same state — shape_index (src/observables.jl:18) and unwrapped_positions
(src/observables.jl:24-29) — evaluated here with numpy from the state the
device holds at that moment. The shape index itself is pinned to the FP64
oracle by test_gpu_parity.py / test_gpu_precision.py and, for a circle, to
2 sqrt(pi) by test_oracle_geometry_energies.py.

Bar: histogram counts identical (the per-body p are the values
fcc_shape_index returns, bins computed with the same FP64 expression);
per-type means within 1e-12 relative (FP64 sums in another order); the
device clock of each sample equal to the simulation's clock. Contours
(src/observables.jl:36-44, fcc_contours) against the oracle within 1e-12
(FP64) / 1e-5 (FP32 and mixed); asynchronous frames equal to the contours
of the state they were requested at.
"""
import sys

import numpy as np
import pytest

import fcc_hip as F

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def nuclei_sim(precision, seed=3, maxmode=6, n=8):
    cells, box = F.hexagonal_monolayer(n, n, packing=0.95, maxmode=maxmode, seed=seed)
    b = F.with_nuclei(cells, ratio=0.4)
    return F.Simulation(b, box, types=(F.BodyType(v0=1.0, Dr=0.2), F.BodyType(Dr=0.1, drag="parent")),
                        energies=(F.OnTypes(F.Tension(0.01), 1),
                                  F.Overlap([[1.0, 1.0], [1.0, 1.0]], related=[[0.0, -2.0], [-2.0, 0.0]])),
                        precision=precision, seed=seed, skin=0.1)


def monolayer_sim(precision, seed=5, n=12, maxmode=8):
    b, box = F.hexagonal_monolayer(n, n, packing=0.95, maxmode=maxmode, seed=seed)
    return F.Simulation(b, box, types=(F.BodyType(v0=2.0, Dr=0.5),), energies=(F.Tension(0.01), F.Overlap(1.0)),
                        precision=precision, seed=seed, skin=0.05)


def host_sample(sim, r0, ntypes, pmin, scale, nbins):
    """The reference's host observables of the current state, per type."""
    p = F.shape_index(sim)
    r = F.unwrapped_positions(sim)
    d2 = np.sum((r - r0) ** 2, axis=0)
    types = sim._bodies.type
    msd, pm, p2 = np.full(ntypes, np.nan), np.full(ntypes, np.nan), np.full(ntypes, np.nan)
    hist = np.zeros((ntypes, nbins + 2), np.uint64)
    for t in range(ntypes):
        sel = types == t + 1
        if sel.any():
            msd[t], pm[t], p2[t] = d2[sel].mean(), p[sel].mean(), (p[sel] ** 2).mean()
        u = np.floor((p[sel] - pmin) * scale)
        b = np.where(p[sel] < pmin, 0, np.where(u < nbins, 1 + np.where(u < nbins, u, 0), nbins + 1))
        np.add.at(hist[t], b.astype(np.int64), 1)
    return msd, pm, p2, hist


@pytest.mark.parametrize("precision,make", [("fp64", nuclei_sim), ("fp32", nuclei_sim), ("fp32", monolayer_sim),
                                            ("mixed_fp16", monolayer_sim), ("mixed_bf16", monolayer_sim)])
def test_samples_equal_host_observables(precision, make):
    say("observer vs host observables", precision, make.__name__)
    sim = make(precision)
    F.run(sim, 20)
    ntypes = len(sim.types)
    nbins, prange = 40, (3.55, 3.75)
    F.observer_start(sim, nbins=nbins, prange=prange, max_samples=8)
    r0 = F.unwrapped_positions(sim)
    scale = nbins / (prange[1] - prange[0])
    want_msd, want_pm, want_p2, clocks = [], [], [], []
    want_hist = np.zeros((ntypes, nbins + 2), np.uint64)
    for s in range(6):
        F.run(sim, 37)
        F.observe(sim)
        m, pm, p2, h = host_sample(sim, r0, ntypes, prange[0], scale, nbins)
        want_msd.append(m)
        want_pm.append(pm)
        want_p2.append(p2)
        want_hist += h
        clocks.append(sim.counters().clock)
    o = F.observations(sim)
    assert o.msd.shape == (6, ntypes) and o.hist.shape == (ntypes, nbins + 2)
    assert np.array_equal(o.clock, clocks)
    assert np.array_equal(o.hist, want_hist)
    assert o.hist.sum() == 6 * sim.N
    # the range catches most bodies, and some fall outside on both sides
    say("  histogram below / inside / above:", o.hist[:, 0].sum(), o.hist[:, 1:-1].sum(), o.hist[:, -1].sum())
    assert o.hist[:, 1:-1].sum() > 0
    np.testing.assert_allclose(o.msd, np.array(want_msd), rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(o.p_mean, np.array(want_pm), rtol=1e-12)
    np.testing.assert_allclose(o.p2_mean, np.array(want_p2), rtol=1e-12)
    assert np.all(np.diff(o.msd[:, 0]) != 0)  # the bodies moved between samples


def test_run_observed_equals_run_and_observe():
    say("run_observed")
    a, b = monolayer_sim("fp32"), monolayer_sim("fp32")
    for s in (a, b):
        F.observer_start(s, nbins=64, prange=(3.5, 3.9), max_samples=16)
    F.run_observed(a, 250, 60)  # 60, 60, 60, 60, 10
    for k in (60, 60, 60, 60, 10):
        F.run(b, k)
        F.observe(b)
    oa, ob = F.observations(a), F.observations(b)
    assert len(oa.clock) == 5 and oa.clock[-1] == a.counters().clock
    for f in ("clock", "msd", "p_mean", "p2_mean", "hist"):
        assert np.array_equal(getattr(oa, f), getattr(ob, f)), f
    assert np.array_equal(a.bodies.x, b.bodies.x)


def test_restart_and_limits():
    say("restart, max_samples, arguments")
    s = monolayer_sim("fp32")
    with pytest.raises(RuntimeError):
        F.observe(s)  # no observation started
    for bad in (dict(nbins=0), dict(nbins=F.api._abi.FCC_OBS_MAX_BINS + 1), dict(max_samples=0),
                dict(prange=(4.0, 4.0)), dict(prange=(3.0, float("inf")))):
        with pytest.raises(F.ArgumentError):
            F.observer_start(s, **bad)
    F.observer_start(s, nbins=8, prange=(3.0, 4.0), max_samples=2)
    F.run_observed(s, 20, 10)
    with pytest.raises(RuntimeError):
        F.observe(s)  # series full
    o = F.observations(s)
    assert len(o.clock) == 2 and o.hist.sum() == 2 * s.N
    # a restart clears the series and the histogram and takes a new origin
    F.observer_start(s, nbins=8, prange=(3.0, 4.0), max_samples=2)
    F.observe(s)
    o = F.observations(s)
    assert len(o.clock) == 1 and o.hist.sum() == s.N
    assert np.all(o.msd == 0.0)


def test_ensemble_observations_equal_replicas_alone():
    say("ensemble observations")
    shapes = [(8, 8), (14, 10), (6, 20)]
    reps = []
    for r, (nx, ny) in enumerate(shapes):
        b, box = F.hexagonal_monolayer(nx, ny, packing=0.95, maxmode=6, seed=40 + r)
        reps.append(F.Replica(b, box, types=(F.BodyType(v0=1.0 + r, Dr=0.3),),
                              energies=(F.Tension(0.01), F.Overlap(1.0)), seed=40 + r, skin=0.05))
    alone = []
    for rp in reps:
        s = F.Simulation(rp.bodies, rp.box, types=rp.types, energies=rp.energies, dt=rp.dt, seed=rp.seed,
                         skin=rp.skin, precision="fp32")
        F.observer_start(s, nbins=32, prange=(3.5, 3.8), max_samples=10)
        F.run_observed(s, 200, 50)
        alone.append(F.observations(s))
    ens = F.Ensemble(reps, precision="fp32")
    ens.observer_start(nbins=32, prange=(3.5, 3.8), max_samples=10)
    ens.run_observed(200, 50)
    for oa, oe in zip(alone, ens.observations()):
        assert len(oe.clock) == 4
        for f in ("clock", "msd", "p_mean", "p2_mean", "hist"):
            assert np.array_equal(getattr(oa, f), getattr(oe, f)), f


@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed_fp16"])
def test_contours_match_oracle(hip, oracle_engine, precision):
    # contours (src/observables.jl:36-44) of deformed bodies, device vs oracle
    say("contours", precision)
    b, box = F.hexagonal_monolayer(6, 6, packing=0.95, maxmode=6, seed=4)
    rng = np.random.default_rng(4)
    b.coef[1:, :] = 0.04 * rng.standard_normal((b.coef.shape[0] - 1, len(b)))
    kw = dict(types=(F.BodyType(),), energies=(F.Tension(0.01),))
    Xa, Ya = F.contours(F.Simulation(b, box, engine=oracle_engine, **kw))
    Xg, Yg = F.contours(F.Simulation(b, box, precision=precision, **kw))
    atol = 1e-12 if precision == "fp64" else 1e-5
    np.testing.assert_allclose(Xg, Xa, rtol=0, atol=atol)
    np.testing.assert_allclose(Yg, Ya, rtol=0, atol=atol)


def test_frames_overlap_the_next_steps():
    say("frames")
    s = monolayer_sim("fp32")
    with pytest.raises(RuntimeError):
        F.frame_read(s)  # nothing requested
    F.run(s, 30)
    X0, Y0 = F.contours(s)
    c0 = s.counters().clock
    F.frame_request(s)
    F.run(s, 50)  # issued while the frame's copy may still be in flight
    X, Y, clock = F.frame_read(s)
    assert clock == c0 and np.array_equal(X, X0) and np.array_equal(Y, Y0)
    for _ in range(3):  # requests reuse the staging buffer once the last copy is done
        F.frame_request(s)
        F.run(s, 10)
    X, Y, clock = F.frame_read(s)
    assert clock == c0 + 50 + 20
    with pytest.raises(RuntimeError):
        F.frame_read(s)  # already read
    F.frame_request(s)
    X1, Y1 = F.contours(s)
    X, Y, clock = F.frame_read(s)
    assert clock == c0 + 80 and np.array_equal(X, X1) and np.array_equal(Y, Y1)
