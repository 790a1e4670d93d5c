"""Bond terms on the device (k_bonds after the forces kernel; fcc_set_bonds)
against the FP64 oracle, which test_oracle_bonds.py pins to the reference's
own bond tests (test/test_parents.jl:83-138) and to finite differences.

Bars: FP64 gradients within rtol 1e-10 and a 20-step FP64 trajectory within
1e-8 of the oracle; FP32 / mixed gradients within 2e-4 of the largest
magnitude; the reference's closed forms at atol 1e-14 (FP64); graph mode
and ensembles bitwise equal to eager / independent runs.
"""
import sys

import numpy as np
import pytest

import fcc_hip as F

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def round_bodies(x, y, type=1, parent=None, R0=None, maxmode=6):
    n = len(x)
    coef = np.zeros((2 * maxmode - 1, n))
    coef[0, :] = 1.0
    return F.Bodies(x, y, np.zeros(n), coef, np.ones(n) if R0 is None else R0, type=type, parent=parent)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_bond_forces_across_the_boundary_and_on_types(hip, precision):
    # test/test_parents.jl:106-123 on the device
    say("bond forces", precision)
    b = round_bodies([9.8, 0.4, 5.0], [3.0, 3.5, 5.0], type=[1, 2, 1])
    box = F.PeriodicBox(10.0, 10.0)
    delta = np.array([9.8 - 0.4 - 10, 3.0 - 3.5])
    cases = [((F.HarmonicBond(0.7),), 0.7),
             ((F.OnTypes(F.HarmonicBond(0.7), 1, 2),), 0.7),
             ((F.OnTypes(F.HarmonicBond(0.7), 1),), 0.0),
             ((F.HarmonicBond(0.7), F.HarmonicBond(0.3)), 1.0)]
    atol = 1e-14 if precision == "fp64" else 1e-6
    for terms, k in cases:
        sim = F.Simulation(b, box, types=(F.BodyType(), F.BodyType()), energies=terms, bonds=[(1, 2)],
                           precision=precision)
        g = F.compute_gradients(sim)
        np.testing.assert_allclose([g.gx[0], g.gy[0]], k * delta, atol=atol)
        np.testing.assert_allclose([g.gx[1], g.gy[1]], -k * delta, atol=atol)
        assert g.gx[2] == 0.0 and g.gy[2] == 0.0


@pytest.mark.parametrize("graph", [False, True])
def test_bonded_nucleus_relaxes_inside_its_cell(hip, graph):
    # test/test_parents.jl:125-138 on the device (FP64), eager and graph mode
    say("nucleus relaxation, graph =", graph)
    mu, k, d0, n, dt = 2.0, 0.5, 0.2, 300, 0.01
    b = round_bodies([5.0, 5.0 + d0, 1.0], [5.0, 5.0, 1.0], type=[1, 2, 1], parent=[1, 1, 3],
                     R0=np.array([1.0, 0.5, 1.0]))
    b.theta[:] = [0.3, -1.0, 2.0]
    sim = F.Simulation(b, F.PeriodicBox(10.0, 10.0), types=(F.BodyType(), F.BodyType(mur=mu, drag="parent")),
                       energies=(F.HarmonicBond(k),), bonds=F.parent_bonds(b), dt=dt, precision="fp64")
    F.run(sim, n) if graph else F.step(sim, n)
    out = sim.bodies
    assert out.x[0] == 5.0 and out.y[0] == 5.0
    assert out.x[1] - 5.0 == pytest.approx(d0 * (1 - mu * k * dt) ** n, rel=1e-12)


def bonded_nuclei(seed=2):
    cells, box = F.hexagonal_monolayer(6, 4, packing=1.0, seed=seed, maxmode=6)
    b = F.with_nuclei(cells, ratio=0.4)
    rng = np.random.default_rng(seed)
    b.coef[1:, :] = 0.04 * rng.standard_normal((b.coef.shape[0] - 1, len(b)))
    N = len(cells)
    b.x[N:] += 0.15 * (rng.random(N) - 0.5)
    b.y[N:] += 0.15 * (rng.random(N) - 0.5)
    energies = (F.Tension(0.1), F.HarmonicBond(0.8, 0.02), F.OnTypes(F.HarmonicBond(0.3), 2),
                F.Overlap([[1.0, 1.0], [1.0, 1.0]], related=[[0.0, -2.0], [-2.0, 0.0]]))
    types = (F.BodyType(v0=0.3, Dr=0.1), F.BodyType(Dr=0.1, drag="parent"))
    return b, box, types, energies


@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed_fp16"])
def test_bonded_nuclei_match_oracle(hip, oracle_engine, precision):
    say("bonded nuclei", precision)
    b, box, types, energies = bonded_nuclei()
    bonds = F.parent_bonds(b)
    a = F.refresh(F.Simulation(b, box, types=types, energies=energies, bonds=bonds, engine=oracle_engine))
    g = F.refresh(F.Simulation(b, box, types=types, energies=energies, bonds=bonds, precision=precision))
    nob = F.refresh(F.Simulation(b, box, types=types, energies=energies, engine=oracle_engine))
    ga, gg = a.grad, g.grad
    assert np.max(np.abs(ga.gx - nob.grad.gx)) > 1e-3  # the bonds contribute
    tol = 1e-10 if precision == "fp64" else 2e-4
    scale = np.max(np.abs(ga.gx))
    np.testing.assert_allclose(gg.gx, ga.gx, rtol=tol, atol=tol * scale)
    np.testing.assert_allclose(gg.gy, ga.gy, rtol=tol, atol=tol * scale)
    np.testing.assert_allclose(gg.gcoef, ga.gcoef, rtol=tol, atol=tol * np.max(np.abs(ga.gcoef)))
    if precision == "fp64":
        F.step(a, 20)
        F.step(g, 20)
        np.testing.assert_allclose(g.bodies.x, a.bodies.x, rtol=1e-8)
        np.testing.assert_allclose(g.bodies.y, a.bodies.y, rtol=1e-8)
        np.testing.assert_allclose(g.bodies.coef, a.bodies.coef, rtol=1e-8, atol=1e-12)


def test_graph_equals_eager_and_bonds_can_change(hip):
    say("bonds: graph == eager, set_bonds between runs")
    b, box, types, energies = bonded_nuclei(seed=4)
    bonds = F.parent_bonds(b)
    e = F.Simulation(b, box, types=types, energies=energies, bonds=bonds, precision="fp32", skin=0.1)
    r = F.Simulation(b, box, types=types, energies=energies, bonds=bonds, precision="fp32", skin=0.1)
    F.step(e, 40)
    F.run(r, 40)
    assert np.array_equal(e.bodies.x, r.bodies.x) and np.array_equal(e.bodies.coef, r.bodies.coef)
    # removing the bonds takes the bond kernel out of the captured step
    for s in (e, r):
        s.set_bonds([])
    F.step(e, 30)
    F.run(r, 30)
    assert np.array_equal(e.bodies.x, r.bodies.x) and np.array_equal(e.bodies.y, r.bodies.y)
    with pytest.raises(F.ArgumentError):
        e.set_bonds([(1, 1)])


def test_ensemble_with_bonds_equals_replicas_alone(hip):
    say("ensemble with bonds")
    reps = []
    for s in (1, 2, 3):
        b, box, types, energies = bonded_nuclei(seed=s)
        reps.append(F.Replica(b, box, types=types, energies=energies, seed=s, skin=0.1,
                              bonds=F.parent_bonds(b) if s != 2 else ()))
    alone = []
    for rp in reps:
        sim = F.Simulation(rp.bodies, rp.box, types=rp.types, energies=rp.energies, dt=rp.dt, seed=rp.seed,
                           skin=rp.skin, precision="fp32", bonds=rp.bonds)
        alone.append(F.run(sim, 80))
    ens = F.Ensemble(reps, precision="fp32")
    ens.run(80)
    for a, m in zip(alone, ens.members):
        assert np.array_equal(a.bodies.x, m.bodies.x) and np.array_equal(a.bodies.coef, m.bodies.coef)
        assert a.counters().clock == m.counters().clock
