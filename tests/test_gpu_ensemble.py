"""Ensembles of independent replicas (fcc_ensemble_*): every replica must end
bitwise where the same simulation run alone ends, as the reference requires of
concurrent simulations, one task and stream each (test/test_cuda.jl:95-104).
The single-simulation path these are compared with is itself pinned to the
FP64 oracle by test_gpu_parity.py."""
import ctypes
import itertools
import sys

import numpy as np
import pytest

import fcc_hip as F

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fast_replica(seed, nx=8, ny=8, v0=2.0, Dr=0.5, eta=0.01, dt=0.005, maxmode=6, skin=0.05, packing=0.95):
    # fast cells and a thin skin: many rebuilds, frozen chunks (test/test_cuda.jl:39-41)
    b, box = F.hexagonal_monolayer(nx, ny, packing=packing, maxmode=maxmode, seed=seed)
    return F.Replica(b, box, types=(F.BodyType(v0=v0, Dr=Dr),),
                     energies=(F.Tension(eta), F.Overlap(1.0)), dt=dt, seed=seed, skin=skin)


def alone(rp, precision, two_lanes=False):
    return F.Simulation(rp.bodies, rp.box, types=rp.types, energies=rp.energies, dt=rp.dt, seed=rp.seed,
                        skin=rp.skin, precision=precision, two_lanes=two_lanes)


def state(s):
    b = s.bodies
    c = s.counters()
    return (b.x, b.y, b.theta, b.coef, b.imgx, b.imgy), (c.step, c.t, c.clock, c.builds)


def assert_same(a, b):
    (sa, ca), (sb, cb) = state(a), state(b)
    for u, v in zip(sa, sb):
        assert np.array_equal(u, v)
    assert ca == cb


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_concurrent_equals_sequential(precision):
    say("ensemble == sequential", precision)
    reps = [fast_replica(seed) for seed in (1, 2, 3)]
    seq = [F.run(alone(rp, precision), 150) for rp in reps]
    ens = F.Ensemble(reps, precision=precision)
    ens.run(150)
    for a, b in zip(seq, ens.members):
        assert_same(a, b)
    assert seq[0].counters().builds > 10
    assert not np.array_equal(ens.members[0].bodies.x, ens.members[1].bodies.x)


@pytest.mark.parametrize("precision,maxmode", [("fp32", 6), ("fp32", 8), ("mixed_fp16", 8),
                                               ("mixed_bf16", 6)])
def test_uneven_sizes_and_parameters(precision, maxmode):
    say("uneven ensemble", precision, maxmode)
    shapes = [(4, 4), (6, 6), (10, 8), (16, 16), (24, 20), (12, 30)]
    reps = [fast_replica(seed=10 + r, nx=nx, ny=ny, v0=0.4 + 0.5 * r, Dr=0.1 * (r + 1), eta=0.005 * (r + 1),
                         dt=0.004 + 0.001 * (r % 2), maxmode=maxmode, skin=0.05 + 0.02 * r,
                         packing=0.9 + 0.02 * r)
            for r, (nx, ny) in enumerate(shapes)]
    seq = [F.run(alone(rp, precision), 100) for rp in reps]
    ens = F.Ensemble(reps, precision=precision)
    ens.run(60)
    ens.run(40)
    for a, b in zip(seq, ens.members):
        assert_same(a, b)


@pytest.mark.parametrize("update_block,xcd_order,lanes,kernels",
                         list(itertools.product([64, 128, 256], [0, 1], [0, 2], [0, 2])))
def test_every_launch_shape_gives_the_same_run(update_block, xcd_order, lanes, kernels):
    # lanes = 2: the two-lane forces kernels for every replica, as replicas
    # created with FCC_FLAG_TWO_LANES would run alone (and then, with
    # kernels = 0, one kernel per step)
    say("launch shape", update_block, xcd_order, lanes, kernels)
    reps = [fast_replica(seed=s, nx=nx, ny=ny) for s, (nx, ny) in enumerate([(6, 6), (20, 14), (8, 30)])]
    seq = [F.run(alone(rp, "fp32", two_lanes=lanes == 2), 80) for rp in reps]
    ens = F.Ensemble(reps, precision="fp32")
    launch = {"update_block": update_block, "xcd_order": xcd_order, "chunk": 0, "lanes": lanes,
              "kernels": kernels}
    ens.launch = launch
    assert ens.launch == launch
    ens.run(80)
    for a, b in zip(seq, ens.members):
        assert_same(a, b)
    # a member handle continues alone with the ensemble's lane choice
    F.step(ens.members[1], 5)
    F.step(seq[1], 5)
    assert_same(seq[1], ens.members[1])


def test_autotune_leaves_the_state_unchanged():
    # replicas created with two lanes per body: the lane choice is not part of
    # the search, and the tuned run continues the untuned one bitwise
    say("autotune")
    reps = [fast_replica(seed=s, nx=nx, ny=ny) for s, (nx, ny) in enumerate([(8, 8), (30, 30), (16, 40)])]
    seq = [F.run(alone(rp, "fp32", two_lanes=True), 50) for rp in reps]
    ens = F.Ensemble(reps, precision="fp32", two_lanes=True)
    ens.run(20)
    launch, ms = ens.autotune(10)
    say("  best", launch, f"{ms:.4f} ms/step")
    assert launch["update_block"] in (64, 128, 256) and launch["xcd_order"] in (0, 1)
    assert launch["lanes"] == 0
    assert ms > 0
    ens.run(30)
    for a, b in zip(seq, ens.members):
        assert_same(a, b)


def test_autotuned_lane_choice_matches_replicas_alone():
    # the autotuner also tries two lanes per body for the small replicas; the
    # replicas then equal replicas run alone with the chosen lane count
    say("autotune with the lane choice")
    reps = [fast_replica(seed=s, nx=nx, ny=ny) for s, (nx, ny) in enumerate([(8, 8), (30, 30), (16, 40)])]
    ens = F.Ensemble(reps, precision="fp32")
    launch, ms = ens.autotune(10)
    say("  best", launch, f"{ms:.4f} ms/step")
    assert launch["lanes"] in (0, 2)
    seq = [F.run(alone(rp, "fp32", two_lanes=launch["lanes"] == 2), 60) for rp in reps]
    ens.run(60)
    for a, b in zip(seq, ens.members):
        assert_same(a, b)


def test_fixed_chunk_and_member_calls():
    say("fixed chunk, member calls")
    reps = [fast_replica(seed=s) for s in (4, 5)]
    seq = [alone(rp, "fp32") for rp in reps]
    for s in seq:
        F.run(s, 30)
        F.step(s, 10)
    ens = F.Ensemble(reps, precision="fp32")
    ens.launch = {"update_block": 128, "xcd_order": 1, "chunk": 7}
    ens.run(30)
    for m in ens.members:
        F.step(m, 10)  # no step budget is left behind
    for a, b in zip(seq, ens.members):
        assert_same(a, b)
        assert np.array_equal(F.shape_index(a), F.shape_index(b))
    # a replica handle belongs to its ensemble
    m = ens.members[0]
    rc = m.engine.destroy(m._h)
    assert rc == F.api._abi.FCC_ERR_STATE
    ens.close()
    assert ens.members[0]._h is None


def test_parent_drag_and_generic_kernel():
    say("nuclei with parent drag; generic kernel (maxmode 5)")
    reps = []
    for r, mm in enumerate([5, 5]):
        cells, box = F.hexagonal_monolayer(6, 6, packing=0.95, maxmode=mm, seed=20 + r)
        b = F.with_nuclei(cells, ratio=0.4)
        reps.append(F.Replica(b, box, types=(F.BodyType(v0=0.5 + r, Dr=0.1), F.BodyType(Dr=0.1, drag="parent")),
                              energies=(F.OnTypes(F.Tension(0.01), 1),
                                        F.Overlap([[1.0, 1.0], [1.0, 1.0]], related=[[0.0, -2.0], [-2.0, 0.0]])),
                              seed=30 + r, skin=0.1))
    seq = [F.run(alone(rp, "fp32"), 60) for rp in reps]
    ens = F.Ensemble(reps, precision="fp32")
    ens.run(60)
    for a, b in zip(seq, ens.members):
        assert_same(a, b)


def test_list_overflow_and_kernel_switch_mid_run():
    # skin 4.1 on a hexagonal lattice: 30 neighbours at the start (below the
    # fast kernels' limit of 32 and the default capacity of 32); compressions
    # push lists past both while the replicas run, which exercises the
    # synchronous regrow and the switch to the generic forces kernel
    say("overflow / kernel switch")
    reps = [fast_replica(seed=s, nx=12, ny=12, v0=3.0, skin=4.1) for s in (1, 2)]
    seq = [F.run(alone(rp, "fp32"), 150) for rp in reps]
    ens = F.Ensemble(reps, precision="fp32")
    ens.run(150)
    caps = [m.neighbors.capacity for m in ens.members]
    top = [int(m.neighbors.count.max()) for m in ens.members]
    say(f"  capacities {caps}, largest lists {top}")
    for a, b in zip(seq, ens.members):
        assert_same(a, b)
        assert a.neighbors.capacity == b.neighbors.capacity
        assert np.array_equal(a.neighbors.idx, b.neighbors.idx)


def test_replicas_must_share_modes_and_precision():
    a = fast_replica(1, maxmode=6)
    b = fast_replica(2, maxmode=8)
    with pytest.raises(F.UnsupportedError):
        F.Ensemble([a, b], precision="fp32")
    with pytest.raises(F.ArgumentError):
        F.Ensemble([], precision="fp32")


@pytest.mark.parametrize("two_kernels", [False, True])
def test_timed_run_equals_run(two_kernels):
    # one-kernel steps time the step kernel as "forces" (update 0)
    say("timed ensemble run", two_kernels)
    reps = [fast_replica(seed=s, nx=12, ny=12) for s in (7, 8)]
    e1 = F.Ensemble(reps, precision="fp32")
    e2 = F.Ensemble(reps, precision="fp32", two_kernels=two_kernels)
    e1.run(70)
    f, u = e2.run_timed(70)
    assert f > 0 and (u > 0 if two_kernels else u == 0)
    for a, b in zip(e1.members, e2.members):
        assert_same(a, b)


@pytest.mark.parametrize("precision", ["fp32", "mixed_fp16"])
def test_large_and_small_replicas_in_one_step(precision):
    # a replica of 132496 bodies (two-lane forces kernel, from 131072 bodies
    # on) next to small ones (four-lane kernel): one ensemble step launches
    # both kernels, the second over the tiles after the first's
    say("large + small replicas", precision)
    mm = 6 if precision == "fp32" else 8
    reps = [fast_replica(seed=41, nx=6, ny=6, maxmode=mm), fast_replica(seed=42, nx=364, ny=364, maxmode=mm),
            fast_replica(seed=43, nx=20, ny=14, maxmode=mm)]
    assert len(reps[1].bodies) >= 131072 > max(len(reps[0].bodies), len(reps[2].bodies))
    seq = [F.run(alone(rp, precision), 30) for rp in reps]
    ens = F.Ensemble(reps, precision=precision)
    ens.run(30)
    for a, b in zip(seq, ens.members):
        assert_same(a, b)
    assert seq[1].counters().builds > 1
