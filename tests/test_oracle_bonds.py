"""Oracle bond terms (HarmonicBond, BondList, parent_bonds) against the
reference's own bond tests (test/test_parents.jl:83-138), central finite
differences of the bond energy E = (k/2)(|Delta| - r0)^2
(src/energies/bonds.jl:8-21) and Newton's third law."""
import math

import numpy as np
import pytest

import fcc_hip as F


def round_bodies(x, y, type=1, parent=None, R0=None, maxmode=6):
    n = len(x)
    coef = np.zeros((2 * maxmode - 1, n))
    coef[0, :] = 1.0
    return F.Bodies(x, y, np.zeros(n), coef, np.ones(n) if R0 is None else R0, type=type, parent=parent)


def center_gradients(sim):
    F.compute_gradients(sim)
    g = sim.grad
    return g.gx, g.gy


def pair_sim(engine, delta, terms, types=(1, 2), L=10.0):
    # body 1 at (5, 5) + delta, body 2 at (5, 5); no other terms
    b = round_bodies([5.0 + delta[0], 5.0], [5.0 + delta[1], 5.0], type=list(types))
    return F.Simulation(b, F.PeriodicBox(L, L), types=(F.BodyType(), F.BodyType()), energies=terms,
                        bonds=[(1, 2)], engine=engine)


def bond_energy(k, r0, d):
    return k / 2 * (math.hypot(d[0], d[1]) - r0) ** 2


@pytest.mark.parametrize("k,r0", [(1.3, 0.0), (1.3, 0.4)])
def test_gradient_matches_finite_differences(oracle_engine, k, r0):
    # test/test_parents.jl:83-87: Delta = (0.3, -0.7)
    d = (0.3, -0.7)
    gx, gy = center_gradients(pair_sim(oracle_engine, d, (F.HarmonicBond(k, r0),)))
    h = 1e-6
    fd = [(bond_energy(k, r0, (d[0] + h, d[1])) - bond_energy(k, r0, (d[0] - h, d[1]))) / (2 * h),
          (bond_energy(k, r0, (d[0], d[1] + h)) - bond_energy(k, r0, (d[0], d[1] - h))) / (2 * h)]
    np.testing.assert_allclose([gx[0], gy[0]], fd, rtol=1e-7)
    # the partner gets the opposite force (its own work item, Delta -> -Delta)
    assert gx[1] == -gx[0] and gy[1] == -gy[0]


def test_coincident_centres(oracle_engine):
    # test/test_parents.jl:88-90
    for r0 in (0.0, 0.5):
        gx, gy = center_gradients(pair_sim(oracle_engine, (0.0, 0.0), (F.HarmonicBond(2.0, r0),)))
        assert gx[0] == 0.0 and gy[0] == 0.0
    assert bond_energy(2.0, 0.5, (0.3, 0.4)) == pytest.approx(0.0)


def test_bond_forces_across_the_boundary_and_on_types(oracle_engine):
    # test/test_parents.jl:106-123: a bonded pair across the periodic boundary,
    # the third body is unbonded
    b = round_bodies([9.8, 0.4, 5.0], [3.0, 3.5, 5.0], type=[1, 2, 1])
    box = F.PeriodicBox(10.0, 10.0)
    delta = np.array([9.8 - 0.4 - 10, 3.0 - 3.5])
    cases = [((F.HarmonicBond(0.7),), 0.7),
             ((F.OnTypes(F.HarmonicBond(0.7), 1, 2),), 0.7),
             ((F.OnTypes(F.HarmonicBond(0.7), 1),), 0.0),  # type 2 excluded
             ((F.HarmonicBond(0.7), F.HarmonicBond(0.3)), 1.0)]
    for terms, k in cases:
        sim = F.Simulation(b, box, types=(F.BodyType(), F.BodyType()), energies=terms, bonds=[(1, 2)],
                           engine=oracle_engine)
        gx, gy = center_gradients(sim)
        np.testing.assert_allclose([gx[0], gy[0]], k * delta, atol=1e-14)
        np.testing.assert_allclose([gx[1], gy[1]], -k * delta, atol=1e-14)
        assert gx[2] == 0.0 and gy[2] == 0.0
    with pytest.raises(F.ArgumentError):
        F.Simulation(b, box, types=(F.BodyType(), F.BodyType()), energies=(F.HarmonicBond(1.0), "notaterm"),
                     engine=oracle_engine)


def test_bond_list_checks(oracle_engine):
    # test/test_parents.jl:93-104
    b = round_bodies([1.0, 4.0, 7.0], [1.0, 4.0, 7.0])
    sim = F.Simulation(b, F.PeriodicBox(10.0, 10.0), types=(F.BodyType(),), energies=(F.HarmonicBond(1.0),),
                       engine=oracle_engine)
    for bad in ([(1, 1)], [(1, 2), (2, 1)], [(1, 4)], [(0, 2)]):
        with pytest.raises(F.ArgumentError):
            sim.set_bonds(bad)
    sim.set_bonds([(1, 2), (3, 1)])
    pb = round_bodies([0.0] * 3, [0.0] * 3, parent=[1, 1, 2])
    assert F.parent_bonds(pb) == [(2, 1), (3, 2)]
    cells, _ = F.hexagonal_monolayer(4, 4)
    nuc = F.with_nuclei(cells, ratio=0.4)
    N = len(cells)
    assert F.parent_bonds(nuc) == [(N + i, i) for i in range(1, N + 1)]


def test_bonded_nucleus_relaxes_inside_its_cell(oracle_engine):
    # test/test_parents.jl:125-138: parent drag makes the bond internal, so the
    # cell stays put and the offset follows explicit Euler of d' = -mu k d
    mu, k, d0, n, dt = 2.0, 0.5, 0.2, 300, 0.01
    b = round_bodies([5.0, 5.0 + d0, 1.0], [5.0, 5.0, 1.0], type=[1, 2, 1], parent=[1, 1, 3],
                     R0=np.array([1.0, 0.5, 1.0]))
    b.theta[:] = [0.3, -1.0, 2.0]
    sim = F.Simulation(b, F.PeriodicBox(10.0, 10.0), types=(F.BodyType(), F.BodyType(mur=mu, drag="parent")),
                       energies=(F.HarmonicBond(k),), bonds=F.parent_bonds(b), dt=dt, engine=oracle_engine)
    F.step(sim, n)
    out = sim.bodies
    assert out.x[0] == 5.0 and out.y[0] == 5.0
    assert out.x[1] - 5.0 == pytest.approx(d0 * (1 - mu * k * dt) ** n, rel=1e-12)


def test_newtons_third_law_with_nuclei(oracle_engine):
    # every bond acts on both ends with opposite forces: the bond forces sum to zero
    cells, box = F.hexagonal_monolayer(6, 6, packing=0.9, maxmode=6, seed=2)
    nuc = F.with_nuclei(cells, ratio=0.4)
    rng = np.random.default_rng(1)
    nuc.x[len(cells):] += 0.2 * (rng.random(len(cells)) - 0.5)
    nuc.y[len(cells):] += 0.2 * (rng.random(len(cells)) - 0.5)
    types = (F.BodyType(), F.BodyType(drag="parent"))
    with_b = F.Simulation(nuc, box, types=types, energies=(F.Tension(0.01), F.HarmonicBond(0.8, 0.05)),
                          bonds=F.parent_bonds(nuc), engine=oracle_engine)
    without = F.Simulation(nuc, box, types=types, energies=(F.Tension(0.01), F.HarmonicBond(0.8, 0.05)),
                           engine=oracle_engine)
    gx1, gy1 = center_gradients(with_b)
    gx0, gy0 = center_gradients(without)
    bx, by = gx1 - gx0, gy1 - gy0
    assert abs(bx.sum()) < 1e-13 and abs(by.sum()) < 1e-13
    assert np.abs(bx).max() > 1e-3
    # each cell feels exactly minus its nucleus' bond force
    N = len(cells)
    np.testing.assert_allclose(bx[:N], -bx[N:], rtol=0, atol=1e-15)
