"""Oracle dynamics vs the reference's fixture (output of the original C++ code),
closed forms and statistics (test/test_dynamics.jl, test/test_parents.jl)."""
import math
import os

import numpy as np
import pytest

import fcc_hip as F
import oracle
from helpers import read_reference

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_N64.txt")


def test_regression_against_original_code(oracle_engine):
    # test/test_dynamics.jl:53-71: N=64, Dr=0, Tension(0.01)+Overlap(1), atol 1e-10
    init, snaps = read_reference(GOLDEN)
    N = init.shape[1]
    Ls = math.sqrt(N * math.pi / (0.95 * 2 * math.sqrt(3)))
    coef = np.zeros((11, N))
    coef[0] = 1.0
    b = F.Bodies(init[0], init[1], init[2], coef, np.ones(N))
    sim = F.Simulation(b, F.PeriodicBox(2 * Ls, math.sqrt(3) * Ls),
                       types=(F.BodyType(v0=0.06, Dr=0.0, muab=0.1),),
                       energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005, engine=oracle_engine)
    assert sorted(snaps) == [1001, 4001]
    for n in sorted(snaps):
        F.step(sim, n - sim.step_count)
        ref = snaps[n]
        pos = F.unwrapped_positions(sim)
        bod = sim.bodies
        np.testing.assert_allclose(pos[0], ref[3], atol=1e-10, rtol=0)
        np.testing.assert_allclose(pos[1], ref[4], atol=1e-10, rtol=0)
        np.testing.assert_allclose(bod.theta, ref[2], atol=1e-10, rtol=0)
        np.testing.assert_allclose(bod.coef[1:], ref[5:15], atol=1e-10, rtol=0)
        # the wrapped coordinates of the fixture too
        np.testing.assert_allclose(bod.x, ref[0], atol=1e-9, rtol=0)
        np.testing.assert_allclose(bod.y, ref[1], atol=1e-9, rtol=0)


def test_hard_area_constraint(oracle_engine):
    # test/test_dynamics.jl:23-31
    b, box = F.hexagonal_monolayer(2, 2)
    b.coef[1:, :] = (0.05 * np.arange(1, 11))[:, None]
    sim = F.Simulation(b, box, types=(F.BodyType(),), ntheta=64, engine=oracle_engine)
    c = sim.bodies.coef
    assert np.all(np.sum(np.hypot(c[1:6], c[6:11]), axis=0) <= math.sqrt(2 / 3) + 1e-12)
    np.testing.assert_allclose(F.areas(sim), np.full(4, math.pi), rtol=1e-12)


def test_isolated_cell_self_propels(oracle_engine):
    # test/test_dynamics.jl:33-43
    coef = np.zeros((11, 1))
    coef[0] = 1
    b = F.Bodies([1.0], [1.0], [0.3], coef, [1.0])
    sim = F.Simulation(b, F.PeriodicBox(5.0, 5.0), types=(F.BodyType(v0=0.5),),
                       energies=(F.Tension(0.1), F.Overlap(1.0)), dt=0.01, engine=oracle_engine)
    F.step(sim, 3000)
    np.testing.assert_allclose(F.unwrapped_positions(sim)[:, 0],
                               [1 + 15 * math.cos(0.3), 1 + 15 * math.sin(0.3)], rtol=1e-10)
    bod = sim.bodies
    assert 0 <= bod.x[0] < 5 and 0 <= bod.y[0] < 5
    assert abs(bod.theta[0] - 0.3) < 1e-12
    assert np.max(np.abs(bod.coef[1:, 0])) < 1e-14
    assert abs(sim.t - 30) < 1e-9


def test_rotational_diffusion(oracle_engine):
    # test/test_dynamics.jl:45-51: <θ²> = 2 Dr t
    N = 4000
    rng = np.random.default_rng(1)
    coef = np.zeros((11, N))
    coef[0] = 1
    b = F.Bodies(rng.random(N), rng.random(N), np.zeros(N), coef, np.ones(N))
    sim = F.Simulation(b, F.PeriodicBox(1.0, 1.0), types=(F.BodyType(Dr=0.1),), dt=0.01, seed=7,
                       engine=oracle_engine)
    F.step(sim, 500)
    th = sim.bodies.theta
    assert abs(np.mean(th ** 2) / (2 * 0.1 * sim.t) - 1) < 0.1


def test_same_seed_same_trajectory(oracle_engine):
    # test/test_random.jl:51-57
    def mk(seed):
        b, box = F.hexagonal_monolayer(4, 4)
        return F.Simulation(b, box, types=(F.BodyType(v0=0.1, Dr=0.5, Dt=0.01),),
                            energies=(F.Tension(0.1), F.Overlap(1.0)), seed=seed, engine=oracle_engine)
    a, b, c = F.step(mk(3), 20), F.step(mk(3), 20), F.step(mk(4), 20)
    assert np.array_equal(a.bodies.theta, b.bodies.theta)
    assert not np.array_equal(a.bodies.theta, c.bodies.theta)


def cell_nucleus_sim(engine, nucleus=None, cell=None, dt=0.01):
    # test/test_parents.jl:18-23
    coef = np.zeros((11, 3))
    coef[0] = 1
    b = F.Bodies([5.0, 5.0, 1.0], [5.0, 5.0, 1.0], [0.3, -1.0, 2.0], coef, [1.0, 0.5, 1.0],
                 type=[1, 2, 1], parent=[1, 1, 3])
    return F.Simulation(b, F.PeriodicBox(10.0, 10.0),
                        types=(cell or F.BodyType(), nucleus or F.BodyType(drag="parent")), dt=dt,
                        engine=engine)


def test_displacements_with_parent_drag(oracle_engine):
    # test/test_parents.jl:50-72
    cell = F.BodyType(v0=0.3, Dt=0.02, mur=0.7)
    nucleus = F.BodyType(v0=0.1, Dt=0.01, mur=2.5, drag="parent")
    sim = cell_nucleus_sim(oracle_engine, nucleus, cell)
    rng = np.random.default_rng(4)
    gx, gy = rng.standard_normal(3), rng.standard_normal(3)
    xi = rng.standard_normal((3, 3))
    b0 = sim.bodies
    oracle.set_center_gradients(sim, gx, gy)
    oracle.integrate_noise(sim, xi)
    b1 = sim.bodies
    dt = sim.dt
    e = lambda i: np.array([math.cos(b0.theta[i]), math.sin(b0.theta[i])])
    G = lambda i: np.array([gx[i], gy[i]])
    noise = lambda i, D: math.sqrt(2 * D * dt) * xi[:2, i]
    dcell = dt * (cell.v0 * e(0) - cell.mur * (G(0) + G(1))) + noise(0, cell.Dt)
    dnuc = dcell + dt * (nucleus.v0 * e(1) - nucleus.mur * G(1)) + noise(1, nucleus.Dt)
    dfree = dt * (cell.v0 * e(2) - cell.mur * G(2)) + noise(2, cell.Dt)
    for i, d in ((0, dcell), (1, dnuc), (2, dfree)):
        np.testing.assert_allclose([b1.x[i] - b0.x[i], b1.y[i] - b0.y[i]], d, rtol=1e-12, atol=1e-15)


def test_nucleus_is_carried_along(oracle_engine):
    # test/test_parents.jl:74-81
    sim = cell_nucleus_sim(oracle_engine, cell=F.BodyType(v0=0.5))
    F.run(sim, 2000)
    X = F.unwrapped_positions(sim)
    np.testing.assert_allclose(X[:, 0], [5 + 10 * math.cos(0.3), 5 + 10 * math.sin(0.3)], rtol=1e-10)
    np.testing.assert_allclose(X[:, 1], X[:, 0], rtol=1e-12)


def test_parent_validation(oracle_engine):
    # test/test_parents.jl:34-48
    coef = np.zeros((11, 2))
    coef[0] = 1
    box = F.PeriodicBox(10.0, 10.0)
    child = F.BodyType(drag="parent")
    with pytest.raises(F.ArgumentError):
        F.BodyType(drag="cytoplasm")
    b = F.Bodies([1.0, 2.0], [1.0, 2.0], np.zeros(2), coef, np.ones(2), type=[1, 2])
    with pytest.raises(F.ArgumentError):
        F.Simulation(b, box, types=(F.BodyType(), child), engine=oracle_engine)
    b = F.Bodies([1.0, 2.0], [1.0, 2.0], np.zeros(2), coef, np.ones(2), type=[2, 2], parent=[2, 1])
    with pytest.raises(F.ArgumentError):
        F.Simulation(b, box, types=(F.BodyType(), child), engine=oracle_engine)
    b = F.Bodies([1.0, 2.0], [1.0, 2.0], np.zeros(2), coef, np.ones(2), parent=[1, 3])
    with pytest.raises(F.ArgumentError):
        F.Simulation(b, box, types=(F.BodyType(),), engine=oracle_engine)
