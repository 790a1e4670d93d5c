"""One-kernel steps (k_step_mfma, k_step_fast2 with two lanes per body,
k_step_mfma4, k_step_fast4 with four): the forces and the update of a step in
one kernel, whose workgroups write the new positions, lab rows and radius
bounds into the other buffer of each pair while other workgroups still read
the current state as neighbours. Every result must equal bitwise that of the
two-kernel step (FCC_FLAG_TWO_KERNELS: the forces kernel, then the update
kernel), whose arithmetic it shares: in graph mode and eagerly, over rebuilds,
frozen chunks, odd step counts, checkpoints, observers and state reads. The
two-kernel step itself is pinned to the FP64 oracle by test_gpu_parity.py and
test_gpu_precision.py; ensembles against single runs are also in
test_gpu_ensemble.py and test_gpu_config4.py."""
import sys

import numpy as np
import pytest

import fcc_hip as F

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def monolayer(precision, two_kernels, n=16, maxmode=8, seed=3, v0=0.6, skin=0.08, types=None, energies=None,
              two_lanes=True):
    b, box = F.hexagonal_monolayer(n, n, packing=0.95, maxmode=maxmode, seed=seed)
    rng = np.random.default_rng(seed)
    K = (b.coef.shape[0] - 1) // 2
    b.coef[1:, :] = rng.uniform(-0.03, 0.03, size=(2 * K, len(b)))
    if types is None:
        types = (F.BodyType(v0=v0, Dr=0.3, Dt=0.002),)
    if energies is None:
        energies = (F.Tension(0.01), F.Overlap(1.0))
    return F.Simulation(b, box, types=types, energies=energies, dt=0.005, skin=skin, seed=seed,
                        precision=precision, two_lanes=two_lanes, two_kernels=two_kernels)


def state(s):
    b = s.bodies
    g = s.grad
    c = s.counters()
    return (b.x, b.y, b.theta, b.coef, b.imgx, b.imgy, g.gx, g.gy, g.gth, g.glab, g.Ebody), \
        (c.step, c.t, c.clock, c.builds)


def assert_same(a, b, what):
    (sa, ca), (sb, cb) = state(a), state(b)
    names = ("x", "y", "theta", "coef", "imgx", "imgy", "gx", "gy", "gth", "glab", "Ebody")
    for name, u, v in zip(names, sa, sb):
        assert np.array_equal(u, v), f"{what}: {name} differs (max {np.max(np.abs(u - v))})"
    assert ca == cb, f"{what}: counters {ca} != {cb}"


@pytest.mark.parametrize("precision", ["fp32", "mixed_fp16", "mixed_bf16"])
@pytest.mark.parametrize("maxmode", [6, 8])
@pytest.mark.parametrize("lanes", [2, 4])
def test_one_kernel_equals_two_kernels(precision, maxmode, lanes):
    say("one kernel == two kernels", precision, maxmode, lanes)
    tl = lanes == 2
    a = monolayer(precision, False, maxmode=maxmode, two_lanes=tl)
    b = monolayer(precision, True, maxmode=maxmode, two_lanes=tl)
    # odd counts: the state ends in either buffer of the pairs
    for n in (137, 1, 64):
        F.run(a, n)
        F.run(b, n)
        assert_same(a, b, f"run {n}")
    F.step(a, 3)
    F.step(b, 3)
    assert_same(a, b, "eager steps")
    assert a.counters().builds > 5
    # a checkpoint taken from either buffer parity resumes alike
    c = monolayer(precision, False, maxmode=maxmode, two_lanes=tl).restore(a.checkpoint())
    for s in (a, b, c):
        F.run(s, 41)
    assert_same(a, b, "after checkpoint")
    assert_same(a, c, "restored")
    # state written between runs (bodies setter) and a rebuilt list
    for s in (a, b):
        bo = s.bodies
        bo.x = bo.x + 0.01
        s.bodies = bo
        F.run(s, 25)
    assert_same(a, b, "after set_bodies")


@pytest.mark.parametrize("precision", ["fp32", "mixed_fp16"])
@pytest.mark.parametrize("lanes", [2, 4])
def test_one_kernel_graph_equals_eager_with_callbacks(precision, lanes):
    say("one kernel graph == eager", precision, lanes)
    a, b = (monolayer(precision, False, two_lanes=lanes == 2) for _ in range(2))
    F.step(a, 90)
    seen = []
    F.run(b, 90, every=7, callback=lambda s: seen.append(s.step_count))
    assert seen == list(range(7, 89, 7)) + [90]
    assert_same(a, b, "graph vs eager")
    a.dt = b.dt = 0.004
    F.step(a, 31)
    F.run(b, 31)
    assert_same(a, b, "after dt change")


@pytest.mark.parametrize("lanes", [2, 4])
def test_one_kernel_two_coupling_classes_and_bending(lanes):
    # two coupling classes (the NCLS = 2 kernels) and Bending in the plain
    # kernel; types 1 and 2 alternate
    say("one kernel, two classes, bending", lanes)
    types = (F.BodyType(v0=0.5, Dr=0.2), F.BodyType(v0=0.2, Dr=0.4, Dt=0.001))
    energies = (F.Tension(0.01), F.OnTypes(F.Bending(0.002), 2),
                F.Overlap([[1.0, 0.5], [0.5, 1.0]]))

    def mk(prec, tk):
        s = monolayer(prec, tk, types=types, energies=energies, two_lanes=lanes == 2)
        bo = s.bodies
        bo.type = (np.arange(len(bo.x)) % 2 + 1).astype(np.int32)
        s.bodies = bo
        return s

    for prec in ("fp32", "mixed_fp16"):
        a, b = mk(prec, False), mk(prec, True)
        F.run(a, 120)
        F.run(b, 120)
        assert_same(a, b, prec)


@pytest.mark.parametrize("lanes", [2, 4])
def test_one_kernel_observer_and_frames(lanes):
    # device observations between runs, a frame requested before a run (its
    # contours are those of the state at the request), contours after
    say("one kernel observer, frames", lanes)
    a, b = (monolayer("mixed_fp16", tk, two_lanes=lanes == 2) for tk in (False, True))
    for s in (a, b):
        F.run(s, 11)
        F.observer_start(s, nbins=64, prange=(3.4, 4.4), max_samples=16)
        F.run_observed(s, 35, 5)
        F.frame_request(s)
        F.run(s, 7)
    oa, ob = F.observations(a), F.observations(b)
    assert len(oa.clock) == 7
    for u, v in ((oa.clock, ob.clock), (oa.msd, ob.msd), (oa.p_mean, ob.p_mean), (oa.hist, ob.hist)):
        assert np.array_equal(u, v)
    fa, fb = F.frame_read(a), F.frame_read(b)
    assert fa[2] == fb[2] == 46
    assert np.array_equal(fa[0], fb[0]) and np.array_equal(fa[1], fb[1])
    for u, v in zip(F.contours(a), F.contours(b)):
        assert np.array_equal(u, v)
    assert_same(a, b, "observer run")


@pytest.mark.parametrize("precision", ["mixed_fp16", "fp32"])
@pytest.mark.parametrize("n,two_lanes", [(512, False), (256, False), (256, True)])
def test_one_kernel_at_full_size(precision, n, two_lanes):
    # the bench workload (N = 262144, default lane choice: two lanes, one
    # kernel) and N = 65536 (four lanes by default, or two), each with many
    # workgroups resident per CU
    say("one kernel large", precision, n, two_lanes)
    def mk(tk):
        b, box = F.hexagonal_monolayer(n, n, packing=1.0, maxmode=8, seed=1234)
        return F.Simulation(b, box, types=(F.BodyType(v0=0.1, Dr=0.01, muab=0.1),),
                            energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005, ntheta=40,
                            precision=precision, seed=1234, skin=0.3, two_kernels=tk,
                            two_lanes=two_lanes)
    a = mk(False)
    F.run(a, 21)
    sa = state(a)
    a.close()
    b = mk(True)
    F.run(b, 21)
    sb = state(b)
    b.close()
    for u, v in zip(sa[0], sb[0]):
        assert np.array_equal(u, v)
    assert sa[1] == sb[1]


def fast_replica(seed, n, v0=2.0, skin=0.05, maxmode=8):
    # fast cells and a thin skin: many rebuilds on the side stream, frozen chunks
    b, box = F.hexagonal_monolayer(n, n, packing=0.95, maxmode=maxmode, seed=seed)
    return F.Replica(b, box, types=(F.BodyType(v0=v0, Dr=0.5, Dt=0.002),),
                     energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005, seed=seed, skin=skin)


def alone(rp, precision, two_kernels):
    return F.Simulation(rp.bodies, rp.box, types=rp.types, energies=rp.energies, dt=rp.dt, seed=rp.seed,
                        skin=rp.skin, precision=precision, two_lanes=True, two_kernels=two_kernels)


@pytest.mark.parametrize("precision", ["fp32", "mixed_fp16"])
def test_one_kernel_ensemble_equals_two_kernel_replicas(precision):
    # the ensemble's one-kernel step (k_gate + k_step_*_ens, per-replica
    # buffer parity) against each replica run alone with two kernels per
    # step; sizes that are not multiples of the 64-body tiles, odd step
    # counts, and member state read between runs
    say("one-kernel ensemble == two-kernel replicas", precision)
    reps = [fast_replica(s, n) for s, n in ((1, 8), (2, 10), (3, 6), (4, 12))]
    ens = F.Ensemble(reps, precision=precision, two_lanes=True)
    solo = [alone(rp, precision, True) for rp in reps]
    for n in (77, 1, 40):
        ens.run(n)
        for s in solo:
            F.run(s, n)
        for a, b in zip(ens.members, solo):
            assert_same(a, b, f"ensemble run {n}")
    assert min(m.counters().builds for m in ens.members) > 5
    ens.run_timed(21)
    for s in solo:
        F.run(s, 21)
    for a, b in zip(ens.members, solo):
        assert_same(a, b, "timed run")
    # the same ensemble forced to two kernels per step ends alike
    ens2 = F.Ensemble(reps, precision=precision, two_lanes=True, two_kernels=True)
    ens2.run(77 + 1 + 40 + 21)
    for a, b in zip(ens.members, ens2.members):
        assert_same(a, b, "two-kernel ensemble")
    ens.close()
    ens2.close()


def test_one_kernel_ensemble_autotune_and_observer():
    # the autotuner times one-kernel and two-kernel candidates (lanes 2 and
    # each replica's own) from a snapshot and restores it; observations of the
    # members between runs
    say("one-kernel ensemble autotune, observer")
    reps = [fast_replica(s, n, v0=0.5, skin=0.1) for s, n in ((5, 8), (6, 12), (7, 10))]
    ens = F.Ensemble(reps, precision="fp32")
    launch, _ = ens.autotune(steps=16)
    ens.run(30)
    solo = [alone(rp, "fp32", True) if launch["lanes"] == 2 else
            F.Simulation(rp.bodies, rp.box, types=rp.types, energies=rp.energies, dt=rp.dt, seed=rp.seed,
                         skin=rp.skin, precision="fp32") for rp in reps]
    for s in solo:
        F.run(s, 30)
    ens.observer_start(nbins=32, prange=(3.4, 4.4), max_samples=8)
    for s in solo:
        F.observer_start(s, nbins=32, prange=(3.4, 4.4), max_samples=8)
    ens.run_observed(45, 9)
    for s in solo:
        F.run_observed(s, 45, 9)
    for a, b in zip(ens.members, solo):
        assert_same(a, b, "after autotune")
        oa, ob = F.observations(a), F.observations(b)
        assert np.array_equal(oa.msd, ob.msd) and np.array_equal(oa.hist, ob.hist)
    ens.close()


def own_lanes(rp, precision, two_kernels):
    return F.Simulation(rp.bodies, rp.box, types=rp.types, energies=rp.energies, dt=rp.dt, seed=rp.seed,
                        skin=rp.skin, precision=precision, two_kernels=two_kernels)


@pytest.mark.parametrize("precision", ["fp32", "mixed_fp16"])
def test_one_kernel_ensemble_with_both_lane_counts(precision):
    # replicas on their own lane choice: the small ones on the four-lane step
    # kernel, one of 132496 bodies on the two-lane one, both launched per
    # step over their tiles; each equals itself alone with two kernels
    say("one-kernel ensemble, four- and two-lane replicas", precision)
    reps = [fast_replica(11, 10), fast_replica(12, 364, v0=0.5, skin=0.2), fast_replica(13, 8)]
    ens = F.Ensemble(reps, precision=precision)
    solo = [own_lanes(rp, precision, True) for rp in reps]
    for n in (45, 3):
        ens.run(n)
        for s in solo:
            F.run(s, n)
        for a, b in zip(ens.members, solo):
            assert_same(a, b, f"ensemble run {n}")
    assert ens.members[0].counters().builds > 3
    ens.close()
    for s in solo:
        s.close()


@pytest.mark.parametrize("two_kernels", [False, True])
@pytest.mark.parametrize("lanes", [2, 4])
def test_short_runs_between_other_calls(two_kernels, lanes):
    # consecutive runs start from the status the run before left (no start
    # snapshot) unless a call in between may have changed it: short graph
    # runs with reads, kernel timing, forced list builds, a dt change,
    # checkpoint round trips, timed runs and eager steps in between, against
    # the same sequence of eager steps
    import ctypes
    say("short runs between other calls", two_kernels, lanes)
    a, b = (monolayer("mixed_fp16", two_kernels, two_lanes=lanes == 2, v0=1.2, skin=0.05) for _ in range(2))
    f, u = ctypes.c_double(), ctypes.c_double()
    for i, n in enumerate((1, 2, 3, 5, 8, 13, 1, 1, 21, 34, 2, 55, 4)):
        F.run(a, n)
        F.step(b, n)
        if i == 1:
            a.counters(), b.counters()
            a.bodies, b.bodies
        elif i == 3:
            assert a.engine.time_kernels(a._h, 3, ctypes.byref(f), ctypes.byref(u)) == 0
        elif i == 4:
            F.update_neighbors(a, True)
            F.update_neighbors(b, True)
        elif i == 6:
            a.dt = b.dt = 0.004
        elif i == 7:
            a.restore(a.checkpoint())
            b.restore(b.checkpoint())
        elif i == 8:
            assert a.engine.run_timed(a._h, 6, ctypes.byref(f), ctypes.byref(u)) == 0
            F.step(b, 6)
        elif i == 10:
            F.step(a, 3)
            F.step(b, 3)
        elif i == 11:
            F.compute_gradients(a)
            F.compute_gradients(b)
        assert_same(a, b, f"after run {i} ({n} steps)")
    assert a.counters().builds > 4
