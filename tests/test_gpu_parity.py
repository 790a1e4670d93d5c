"""Device path (libfcc_hip.so through the C ABI) vs the FP64 oracle:
bit-exact Philox words and cell / neighbour lists, FP64 gradients and
trajectories to the tolerances of test/test_cuda.jl, the N = 64 fixture of the
original code, and graph mode == eager bitwise."""
import math
import os
import sys

import numpy as np
import pytest

import fcc_hip as F
import oracle
from helpers import read_reference

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_N64.txt")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_philox_words_bit_exact(hip):
    say("philox words")
    import ctypes
    rng = np.random.default_rng(0)
    n = 4096
    ctr = rng.integers(0, 2 ** 32, size=(n, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2 ** 32, size=(n, 2), dtype=np.uint64).astype(np.uint32)
    ctr[:3] = [[0, 0, 0, 0], [0xFFFFFFFF] * 4, [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344]]
    key[:3] = [[0, 0], [0xFFFFFFFF] * 2, [0xA4093822, 0x299F31D0]]
    out = np.zeros((n, 4), dtype=np.uint32)
    u32 = lambda a: np.ascontiguousarray(a).ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    assert hip.philox4x32(u32(ctr), u32(key), n, u32(out)) == 0
    assert tuple(out[0]) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    assert tuple(out[1]) == (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)
    assert tuple(out[2]) == (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)
    for q in range(3, 200):
        assert tuple(out[q]) == oracle.philox4x32(tuple(ctr[q]), tuple(key[q]))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noise_matches_oracle(hip, oracle_engine, prec):
    # test/test_cuda.jl:22-29: rtol 4 eps(T)
    say("noise", prec)
    g = F.philox_noise(100_000, 9, 12, prec, engine=hip)
    c = F.philox_noise(100_000, 9, 12, prec, engine=oracle_engine)
    eps = np.finfo(np.float64 if prec == "fp64" else np.float32).eps
    np.testing.assert_allclose(g, c, rtol=4 * eps, atol=4 * eps)
    big = F.philox_noise(3, 9, 12 + (1 << 32), prec, engine=hip)
    assert not np.array_equal(big, g[:, :3])


def jittered(nx, ny, packing=0.98, seed=3, amp=0.4, modes=6):
    b, box = F.hexagonal_monolayer(nx, ny, packing=packing, seed=2, maxmode=modes)
    rng = np.random.default_rng(seed)
    N = len(b)
    b.x = np.mod(b.x + amp * rng.standard_normal(N), box.Lx)
    b.y = np.mod(b.y + amp * rng.standard_normal(N), box.Ly)
    b.coef[1:, :] = 0.05 * rng.standard_normal((b.coef.shape[0] - 1, N))
    return b, box


@pytest.mark.parametrize("size,prec", [((8, 8), "fp64"), ((32, 32), "fp64"), ((32, 32), "fp32"),
                                       ((512, 512), "fp32")])
def test_lists_bit_exact(size, prec):
    say("lists", size, prec)
    modes = 8 if size[0] >= 512 else 6
    b, box = jittered(*size, modes=modes)
    sim = F.Simulation(b, box, types=(F.BodyType(),), energies=(F.Overlap(1.0),), precision=prec)
    F.update_neighbors(sim, force=True)
    bod, geo = sim.bodies, sim.geom
    cutoff = 2 * 1.0 * (1 + math.sqrt(2 / 3)) + 0.3
    ref = oracle.build_lists(bod.x, bod.y, geo.rmax, box.Lx, box.Ly, cutoff, 0.3,
                             capacity=sim.neighbors.capacity)
    cl, nl = sim.cells, sim.neighbors
    assert (cl.ncx, cl.ncy) == (ref["ncx"], ref["ncy"])
    for k in ("cellid", "perm", "start", "stop"):
        assert np.array_equal(getattr(cl, k), ref[k]), k
    assert np.array_equal(nl.count, ref["count"])
    assert np.array_equal(nl.idx, ref["idx"])


def layout_bodies(nx=5, ny=4):
    # test/test_layout.jl:21-29 (nucleated, deformed), numpy generators
    cells, box = F.hexagonal_monolayer(nx, ny, packing=0.97, seed=2)
    b = F.with_nuclei(cells)
    rng = np.random.default_rng(5)
    b.coef[1:, :] = 0.04 * rng.standard_normal((b.coef.shape[0] - 1, len(b)))
    return b, box


LAYOUT_TYPES = (F.BodyType(v0=0.2, Dr=0.1), F.BodyType(Dr=0.1, drag="parent"))
LAYOUT_ENERGIES = (F.Tension(0.1), F.OnTypes(F.PerimeterElasticity(0.2, 3.7), 2),
                   F.OnTypes(F.Bending(0.01), 2),
                   F.Overlap([[1.0, 1.0], [1.0, 1.0]], related=[[1.0, -2.0], [-2.0, 1.0]]))


def layout_sim(engine, ntheta=40, precision="fp64", v0=0.2, skin=0.3, seed=7, **kw):
    b, box = layout_bodies()
    types = (F.BodyType(v0=v0, Dr=0.1), F.BodyType(Dr=0.1, drag="parent"))
    return F.Simulation(b, box, types=types, energies=LAYOUT_ENERGIES, ntheta=ntheta, precision=precision,
                        seed=seed, skin=skin, debug=True, engine=engine, **kw)


# two_lanes: the two-lane forces kernels (which the library otherwise takes
# from 131072 bodies on) on the same small systems
@pytest.mark.parametrize("ntheta,two_lanes", [(40, False), (40, True), (37, False)])
def test_gradients_match_oracle_fp64(hip, oracle_engine, ntheta, two_lanes):
    # test/test_cuda.jl:10-20: rtol 1e-10, atol 1e-10
    say("gradients fp64 ntheta", ntheta, "two lanes" if two_lanes else "")
    a = F.refresh(layout_sim(oracle_engine, ntheta))
    b = F.refresh(layout_sim(hip, ntheta, two_lanes=two_lanes))
    ga, gb = a.grad, b.grad
    assert np.count_nonzero(ga.P > 0) > 100
    np.testing.assert_allclose(b.geom.R, a.geom.R, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(b.geom.ddR, a.geom.ddR, rtol=1e-12, atol=1e-12)
    for f in ("P", "GR", "GD", "GDD", "gx", "gy", "gth", "gcoef", "Ebody"):
        np.testing.assert_allclose(getattr(gb, f), getattr(ga, f), rtol=1e-10, atol=1e-10, err_msg=f)


def test_short_trajectory_matches_oracle_fp64(hip, oracle_engine):
    # test/test_cuda.jl:31-39: 30 steps, rtol 1e-8, same rebuild count
    say("trajectory fp64")
    a, b = layout_sim(oracle_engine), layout_sim(hip)
    F.step(a, 30)
    F.step(b, 30)
    ba, bb = a.bodies, b.bodies
    np.testing.assert_allclose(bb.x, ba.x, rtol=1e-8)
    np.testing.assert_allclose(bb.theta, ba.theta, rtol=1e-8)
    np.testing.assert_allclose(bb.coef, ba.coef, rtol=1e-8, atol=1e-12)
    assert b.counters().builds == a.counters().builds
    assert b.counters().clock == a.counters().clock == 30


def test_regression_fixture_on_device_fp64(hip):
    # test/test_dynamics.jl:53-71 on the device, atol 1e-10
    say("fixture fp64")
    init, snaps = read_reference(GOLDEN)
    N = init.shape[1]
    Ls = math.sqrt(N * math.pi / (0.95 * 2 * math.sqrt(3)))
    coef = np.zeros((11, N))
    coef[0] = 1.0
    b = F.Bodies(init[0], init[1], init[2], coef, np.ones(N))
    sim = F.Simulation(b, F.PeriodicBox(2 * Ls, math.sqrt(3) * Ls),
                       types=(F.BodyType(v0=0.06, Dr=0.0, muab=0.1),),
                       energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005)
    for n in sorted(snaps):
        F.step(sim, n - sim.step_count)
        ref = snaps[n]
        pos = F.unwrapped_positions(sim)
        bod = sim.bodies
        np.testing.assert_allclose(pos[0], ref[3], atol=1e-10, rtol=0)
        np.testing.assert_allclose(pos[1], ref[4], atol=1e-10, rtol=0)
        np.testing.assert_allclose(bod.theta, ref[2], atol=1e-10, rtol=0)
        np.testing.assert_allclose(bod.coef[1:], ref[5:15], atol=1e-10, rtol=0)


def soft_fluid(precision="fp64", v0=2.0, skin=0.05, seed=1, n=8, ntheta=40, two_kernels=False):
    b, box = F.hexagonal_monolayer(n, n, packing=0.95)
    return F.Simulation(b, box, types=(F.BodyType(v0=v0, Dr=0.5, Dt=0.01),),
                        energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005, skin=skin, seed=seed,
                        precision=precision, ntheta=ntheta, two_kernels=two_kernels)


def state(s):
    b = s.bodies
    g = s.grad
    c = s.counters()
    return (b.x, b.y, b.theta, b.coef, b.imgx, g.gx, g.glab), (c.t, c.step, c.builds, c.clock)


def same(a, b):
    sa, ca = state(a)
    sb, cb = state(b)
    return all(np.array_equal(u, v) for u, v in zip(sa, sb)) and ca == cb


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_graph_mode_equals_eager(prec):
    # test/test_cuda.jl:41-93: many rebuilds, frozen chunks, callbacks, dt change
    say("graph vs eager", prec)
    a, b = soft_fluid(prec), soft_fluid(prec)
    F.step(a, 200)
    F.run(b, 200)
    assert a.counters().builds > 10
    assert same(a, b)
    a, b = soft_fluid(prec), soft_fluid(prec)
    seen = []
    F.step(a, 100)
    F.run(b, 100, every=7, callback=lambda s: seen.append(s.step_count))
    assert seen == list(range(7, 99, 7)) + [100]
    assert same(a, b)
    a.dt = b.dt = 0.004
    F.step(a, 30)
    F.run(b, 30)
    assert same(a, b)
    for s in (a, b):
        s.set_neighbor_capacity(1)
    F.step(a, 30)
    F.run(b, 30)
    assert b.neighbors.capacity > 1
    assert same(a, b)


def test_body_terms_only_graph(hip):
    say("body terms only graph")
    def mk():
        b, box = F.hexagonal_monolayer(4, 4)
        return F.Simulation(b, box, types=(F.BodyType(v0=0.3, Dr=0.5),),
                            energies=(F.PerimeterElasticity(0.2, 3.7),))
    assert same(F.step(mk(), 50), F.run(mk(), 50))


@pytest.mark.parametrize("two_kernels", [False, True])
def test_timed_run_equals_run(hip, two_kernels):
    # the instrumented graph run (bench timing) changes nothing in the results;
    # a one-kernel step is timed as "forces" (update 0)
    import ctypes
    say("timed run == run", two_kernels)
    a, b = soft_fluid("fp32"), soft_fluid("fp32", two_kernels=two_kernels)
    F.run(a, 150)
    f, u = ctypes.c_double(), ctypes.c_double()
    assert hip.run_timed(b._h, 150, ctypes.byref(f), ctypes.byref(u)) == 0
    assert f.value > 0 and (u.value > 0 if two_kernels else u.value == 0)
    assert same(a, b)


@pytest.mark.parametrize("precision", ["fp32", "mixed_fp16"])
def test_kernel_timing_leaves_state_unchanged(hip, precision):
    # fcc_time_kernels saves and restores the state around its launches
    import ctypes
    say("kernel timing", precision)
    a, b = soft_fluid(precision), soft_fluid(precision, two_kernels=True)
    F.run(a, 60)
    F.run(b, 60)
    f, u = ctypes.c_double(), ctypes.c_double()
    assert hip.time_kernels(b._h, 7, ctypes.byref(f), ctypes.byref(u)) == 0
    assert f.value > 0 and u.value > 0
    # and a simulation on one-kernel steps
    assert hip.time_kernels(a._h, 7, ctypes.byref(f), ctypes.byref(u)) == 0
    assert f.value > 0 and np.isfinite(u.value)
    F.run(a, 60)
    F.run(b, 60)
    assert same(a, b)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_generic_kernel_matches_oracle(hip, oracle_engine, precision):
    # K = 4 (maxmode 5) and three Overlap coupling values take the generic
    # forces kernel (per-node float contact density in LDS)
    say("generic kernel", precision)
    cells, box = F.hexagonal_monolayer(6, 4, packing=0.97, seed=2, maxmode=5)
    b = F.with_nuclei(cells)
    rng = np.random.default_rng(8)
    b.coef[1:, :] = 0.04 * rng.standard_normal((b.coef.shape[0] - 1, len(b)))
    energies = (F.Tension(0.1), F.Overlap([[1.0, 0.5], [0.5, 1.0]], related=[[1.0, -2.0], [-2.0, 1.0]]))
    types = (F.BodyType(v0=0.2, Dr=0.1), F.BodyType(Dr=0.1, drag="parent"))
    a = F.refresh(F.Simulation(b, box, types=types, energies=energies, debug=True, engine=oracle_engine))
    g = F.refresh(F.Simulation(b, box, types=types, energies=energies, debug=True, precision=precision))
    tol = 1e-10 if precision == "fp64" else 2e-4
    ga, gg = a.grad, g.grad
    assert np.count_nonzero(ga.P != 0) > 50
    if precision == "fp64":
        np.testing.assert_allclose(gg.P, ga.P, rtol=tol, atol=tol)
    scale = np.max(np.abs(ga.gx))
    np.testing.assert_allclose(gg.gx, ga.gx, rtol=tol, atol=tol * scale)
    np.testing.assert_allclose(gg.gcoef, ga.gcoef, rtol=tol, atol=tol * np.max(np.abs(ga.gcoef)))


@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed_fp16"])
def test_linear_repulsion_matches_oracle(hip, oracle_engine, precision):
    # LinearRepulsion g(d) = k·d for d > 0 (src/energies/contact.jl:30-41),
    # restricted to cells, with Overlap on all bodies: the generic forces
    # kernel accumulates the depth-dependent density per node
    say("linear repulsion", precision)
    cells, box = F.hexagonal_monolayer(6, 4, packing=1.05, seed=2, maxmode=6)
    b = F.with_nuclei(cells, ratio=0.4)
    rng = np.random.default_rng(9)
    b.coef[1:, :] = 0.04 * rng.standard_normal((b.coef.shape[0] - 1, len(b)))
    energies = (F.Tension(0.1), F.OnTypes(F.LinearRepulsion(3.0), 1),
                F.Overlap([[1.0, 1.0], [1.0, 1.0]], related=[[0.0, -2.0], [-2.0, 0.0]]))
    types = (F.BodyType(v0=0.2, Dr=0.1), F.BodyType(Dr=0.1, drag="parent"))
    a = F.refresh(F.Simulation(b, box, types=types, energies=energies, debug=True, engine=oracle_engine))
    g = F.refresh(F.Simulation(b, box, types=types, energies=energies, debug=True, precision=precision))
    tol = 1e-10 if precision == "fp64" else 2e-4
    ga, gg = a.grad, g.grad
    depth = ga.P[(ga.P != 0) & (ga.P != 1.0) & (ga.P != -2.0)]
    assert depth.size > 20  # nodes where the depth-dependent term contributes
    if precision == "fp64":
        np.testing.assert_allclose(gg.P, ga.P, rtol=tol, atol=tol)
    scale = np.max(np.abs(ga.gx))
    np.testing.assert_allclose(gg.gx, ga.gx, rtol=tol, atol=tol * scale)
    np.testing.assert_allclose(gg.gy, ga.gy, rtol=tol, atol=tol * scale)
    np.testing.assert_allclose(gg.gth, ga.gth, rtol=tol, atol=tol * np.max(np.abs(ga.gth)))
    np.testing.assert_allclose(gg.gcoef, ga.gcoef, rtol=tol, atol=tol * np.max(np.abs(ga.gcoef)))
    if precision == "fp64":
        F.step(a, 20)
        F.step(g, 20)
        np.testing.assert_allclose(g.bodies.x, a.bodies.x, rtol=1e-8)
        np.testing.assert_allclose(g.bodies.coef, a.bodies.coef, rtol=1e-8, atol=1e-12)


def clusters(k=12, sites=4, spacing=5.0, jitter=0.15, seed=9, maxmode=6):
    """Clusters of k deformed bodies around the sites of a square lattice
    (sites farther apart than the cutoff): a node is inside up to k - 1 other
    bodies, so the contact counts reach every bit plane of the forces
    kernels' counters (the planes of counts 4 and more, and their carries,
    are only touched when some lane of a wave has such a count)."""
    rng = np.random.default_rng(seed)
    L = sites * spacing
    sx, sy = np.meshgrid((np.arange(sites) + 0.5) * spacing, (np.arange(sites) + 0.5) * spacing)
    n = sites * sites * k
    x = np.repeat(sx.ravel(), k) + jitter * rng.standard_normal(n)
    y = np.repeat(sy.ravel(), k) + jitter * rng.standard_normal(n)
    nc = 2 * (maxmode - 1) + 1
    coef = np.zeros((nc, n), order="F")
    coef[0] = 1.0
    coef[1:, :] = 0.04 * rng.standard_normal((nc - 1, n))
    return F.Bodies(x, y, rng.uniform(0, 2 * np.pi, n), coef, np.ones(n)), F.PeriodicBox(L, L)


@pytest.mark.parametrize("two_lanes", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed_fp16", "mixed_bf16"])
def test_dense_clusters_contact_counts(hip, oracle_engine, precision, two_lanes):
    say("dense clusters", precision, "two lanes" if two_lanes else "")
    b, box = clusters()
    kw = dict(types=(F.BodyType(),), energies=(F.Tension(0.01), F.Overlap(1.0)), debug=True)
    a = F.refresh(F.Simulation(b, box, engine=oracle_engine, **kw))
    g = F.refresh(F.Simulation(b, box, precision=precision, two_lanes=two_lanes, **kw))
    ga, gg = a.grad, g.grad
    assert ga.P.max() >= 8 and np.count_nonzero(ga.P >= 4) > 1000
    if precision == "fp64":
        for f in ("P", "GR", "GD", "gx", "gy", "gth", "gcoef", "Ebody"):
            np.testing.assert_allclose(getattr(gg, f), getattr(ga, f), rtol=1e-10, atol=1e-10, err_msg=f)
        return
    # FP32 contact: a node within rounding of a neighbour's contour may flip
    differ = gg.P != ga.P
    say("  flips", int(differ.sum()), "of", int(np.count_nonzero(ga.P)), "contact nodes")
    assert differ.sum() <= 2e-3 * np.count_nonzero(ga.P)
    assert gg.P.max() == ga.P.max()
    clean = ~differ.any(axis=0)
    assert clean.sum() >= 0.9 * len(clean)
    for f in ("gx", "gy", "gth"):
        ref, got = getattr(ga, f)[clean], getattr(gg, f)[clean]
        np.testing.assert_allclose(got, ref, rtol=2e-4, atol=2e-4 * np.max(np.abs(ref)), err_msg=f)
    ref, got = ga.gcoef[:, clean], gg.gcoef[:, clean]
    np.testing.assert_allclose(got, ref, rtol=2e-4, atol=2e-4 * np.max(np.abs(ref)))


def edge_bodies():
    """A dilute deformed monolayer, shifted off the lattice origin, with the
    bodies nearest each box edge moved exactly onto x = 0, y = 0 or just below
    x = L, y = L and oriented out of or along the box."""
    b, box = F.hexagonal_monolayer(6, 6, packing=0.6, seed=4)
    rng = np.random.default_rng(8)
    b.coef[1:, :] = 0.02 * rng.standard_normal((b.coef.shape[0] - 1, len(b)))
    b.x = np.mod(b.x + 0.137, box.Lx)
    b.y = np.mod(b.y + 0.093, box.Ly)
    top_x, top_y = np.nextafter(box.Lx, 0.0), np.nextafter(box.Ly, 0.0)
    ox, oy = np.argsort(b.x), np.argsort(b.y)
    for i, th in zip(ox[:3], (math.pi, 0.0, 1.25 * math.pi)):
        b.x[i], b.theta[i] = 0.0, th
    for i, th in zip(ox[-3:], (0.0, math.pi, 0.25 * math.pi)):
        b.x[i], b.theta[i] = top_x, th
    for i, th in zip(oy[:2], (-math.pi / 2, math.pi / 2)):
        b.y[i], b.theta[i] = 0.0, th
    for i, th in zip(oy[-2:], (math.pi / 2, -math.pi / 2)):
        b.y[i], b.theta[i] = top_y, th
    return b, box


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_wrap_at_the_box_edges_matches_oracle(hip, oracle_engine, precision):
    # box.jl:21-30 through the update's FP64 wrap, for bodies exactly on the
    # box edges (the wrap's division-free case 0 < x < L and the full one)
    say("wrap edges", precision)
    b, box = edge_bodies()
    kw = dict(types=(F.BodyType(v0=2.0),), energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005, skin=0.3)
    dev = F.Simulation(b, box, precision=precision, engine=hip, **kw)
    ref = F.Simulation(b, box, engine=oracle_engine, **kw)
    tol = 1e-12 if precision == "fp64" else 1e-7
    for n in (1, 2, 3):
        F.step(dev, 1)
        F.step(ref, 1)
        bd, br = dev.bodies, ref.bodies
        assert np.array_equal(bd.imgx, br.imgx) and np.array_equal(bd.imgy, br.imgy), n
        assert np.all((bd.x >= 0.0) & (bd.x < box.Lx) & (bd.y >= 0.0) & (bd.y < box.Ly))
        np.testing.assert_allclose(bd.x, br.x, rtol=0, atol=tol * box.Lx)
        np.testing.assert_allclose(bd.y, br.y, rtol=0, atol=tol * box.Ly)
    assert np.count_nonzero(bd.imgx) + np.count_nonzero(bd.imgy) >= 3


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_four_body_types_match_oracle(hip, oracle_engine, precision):
    # BodyType parameters per type (src/bodies.jl:30-45) for the maximum number
    # of types: each body's v0, Dr, Dt, μr, μθ, μab come from its own type
    say("four types", precision)
    b, box = F.hexagonal_monolayer(8, 8, packing=0.9, seed=6)
    rng = np.random.default_rng(21)
    b.coef[1:, :] = 0.03 * rng.standard_normal((b.coef.shape[0] - 1, len(b)))
    b.type = (np.arange(len(b)) % 4 + 1).astype(np.int32)
    types = (F.BodyType(v0=0.3, Dr=0.02, Dt=0.0, mur=1.0, mutheta=1.0, muab=0.1),
             F.BodyType(v0=0.0, Dr=0.0, Dt=0.01, mur=0.5, mutheta=2.0, muab=0.05),
             F.BodyType(v0=0.1, Dr=0.05, Dt=0.02, mur=2.0, mutheta=0.5, muab=0.2),
             F.BodyType(v0=0.2, Dr=0.0, Dt=0.0, mur=1.5, mutheta=1.0, muab=0.0))
    kw = dict(types=types, energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005, skin=0.3, seed=5)
    dev = F.Simulation(b, box, precision=precision, engine=hip, **kw)
    ref = F.Simulation(b, box, engine=oracle_engine, **kw)
    F.step(dev, 5)
    F.step(ref, 5)
    bd, br = dev.bodies, ref.bodies
    tol = 1e-10 if precision == "fp64" else 2e-5
    np.testing.assert_allclose(bd.x, br.x, rtol=tol, atol=tol)
    np.testing.assert_allclose(bd.y, br.y, rtol=tol, atol=tol)
    np.testing.assert_allclose(bd.theta, br.theta, rtol=tol, atol=tol)
    np.testing.assert_allclose(bd.coef, br.coef, rtol=0, atol=tol)
    # the types differ in every parameter: the same run with all bodies of type 1 differs
    b1 = F.Bodies(b.x, b.y, b.theta, b.coef, b.R0)
    one = F.Simulation(b1, box, precision=precision, engine=hip, **kw)
    F.step(one, 5)
    assert not np.allclose(one.bodies.theta, bd.theta, rtol=0, atol=1e-6)
