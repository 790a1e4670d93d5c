"""Oracle contact, cell and neighbour lists vs brute force, the exact two-disc
force and finite differences of the field energy (test/test_contact.jl)."""
import math

import numpy as np

import fcc_hip as F
import oracle
from conftest import deformed_coef
from helpers import brute_force_pressure, field_energy


def overlap_fn(J, Jrel=None, lin=None):
    J = np.atleast_2d(J)
    Jrel = J if Jrel is None else np.atleast_2d(Jrel)

    def g(d, ti, tj, rel):
        tab = Jrel if rel else J
        v = tab[ti - 1, tj - 1] if tab.shape[0] > 1 else tab[0, 0]
        out = np.where(d >= 0, v, 0.0)
        if lin is not None:
            out = out + np.where(d > 0, lin * d, 0.0)
        return out
    return g


def jittered_monolayer():
    # test/test_contact.jl:70-76 with numpy generators
    b, box = F.hexagonal_monolayer(8, 8, packing=0.98, seed=2)
    rng = np.random.default_rng(3)
    b.x = np.mod(b.x + 0.4 * rng.standard_normal(64), box.Lx)
    b.y = np.mod(b.y + 0.4 * rng.standard_normal(64), box.Ly)
    b.coef[1:, :] = 0.05 * rng.standard_normal((10, 64))
    return b, box


def test_cell_list_matches_brute_force(oracle_engine):
    b, box = jittered_monolayer()
    sim = F.Simulation(b, box, types=(F.BodyType(),), energies=(F.Overlap(1.0), F.LinearRepulsion(0.3)),
                       engine=oracle_engine)
    F.refresh(sim)
    cl = sim.cells
    assert cl.ncx >= 3 and cl.ncy >= 3
    g, geo, bod = sim.grad, sim.geom, sim.bodies
    assert np.count_nonzero(g.P > 0) > 100
    ref = brute_force_pressure(bod, geo.R, geo.lab, box, overlap_fn(1.0, lin=0.3), 40)
    np.testing.assert_allclose(g.P, ref, rtol=1e-12, atol=1e-12)
    # capacity growth (test/test_contact.jl:81-85)
    sim.set_neighbor_capacity(2)
    F.refresh(sim)
    nl = sim.neighbors
    assert nl.capacity >= nl.count.max() > 2
    np.testing.assert_allclose(sim.grad.P, ref, rtol=1e-12, atol=1e-12)


def test_cell_list_structure(oracle_engine):
    # perm sorted by cell, stable (ties ascending), start/stop delimit cells
    b, box = jittered_monolayer()
    sim = F.Simulation(b, box, types=(F.BodyType(),), energies=(F.Overlap(1.0),), engine=oracle_engine)
    F.update_neighbors(sim, force=True)
    cl = sim.cells
    keys = cl.cellid[cl.perm - 1]
    assert np.all(np.diff(keys) >= 0)
    assert np.array_equal(cl.perm, np.lexsort((np.arange(64), cl.cellid)) + 1)
    for c in range(cl.ncx * cl.ncy):
        members = np.flatnonzero(cl.cellid == c + 1)
        if members.size == 0:
            assert (cl.start[c], cl.stop[c]) == (1, 0)
        else:
            assert np.array_equal(np.sort(cl.perm[cl.start[c] - 1:cl.stop[c]] - 1), members)


def two_body_sim(engine, terms, d=1.5, R0=(1.0, 1.0), ntheta=40, deformed=False, theta=(0.3, -1.1),
                 type=(1, 1), parent=(1, 2), ntypes=1):
    # test/test_contact.jl:25-35
    coef = np.zeros((11, 2))
    coef[0] = 1
    if deformed:
        coef[:, 0] = deformed_coef(5)
        coef[:, 1] = deformed_coef(6)
    b = F.Bodies([5.0, 5.0 + d], [5.0, 5.0], list(theta), coef, list(R0), type=list(type),
                 parent=list(parent))
    sim = F.Simulation(b, F.PeriodicBox(10.0, 10.0), types=tuple(F.BodyType() for _ in range(ntypes)),
                       energies=terms, ntheta=ntheta, engine=engine)
    return F.refresh(sim)


def test_two_discs_exact_force(oracle_engine):
    # test/test_contact.jl:88-96: dA/dd = -√(4 - d²)
    d = 1.5
    sim = two_body_sim(oracle_engine, (F.Overlap(1.0),), d=d, ntheta=8000)
    g = sim.grad
    assert abs(g.gx[0] / math.sqrt(4 - d * d) - 1) < 1e-3
    assert abs(g.gx[1] / -math.sqrt(4 - d * d) - 1) < 1e-3
    assert abs(g.gy[0]) < 1e-6


def test_fd_field_energy_linear_repulsion(oracle_engine):
    # test/test_contact.jl:98-120
    sim = two_body_sim(oracle_engine, (F.LinearRepulsion(1.0),), d=1.6, deformed=True, ntheta=200)
    g = sim.grad
    gx, gth, gcoef = g.gx[0], g.gth[0], g.gcoef[1:, 0]
    b0 = sim.bodies
    fn = overlap_fn(0.0, lin=1.0)

    def E_at(dx=0.0, dth=0.0, dc=None):
        b = b0.copy()
        if dc is not None:
            b.coef[1:, 0] += dc
        b.x[0] = 5.0 + dx
        b.theta[0] = 0.3 + dth
        sim.bodies = b
        F.refresh(sim)
        geo = sim.geom
        return field_energy(sim.bodies, geo.R, geo.lab, 200, fn)

    h = 1e-5
    assert abs((E_at(dx=h) - E_at(dx=-h)) / (2 * h) / gx - 1) < 2e-3
    assert abs((E_at(dth=h) - E_at(dth=-h)) / (2 * h) / gth - 1) < 1e-4
    for r in (0, 3, 7):
        e = np.zeros(10)
        e[r] = h
        fd = (E_at(dc=e) - E_at(dc=-e)) / (2 * h)
        assert abs(fd - gcoef[r]) <= 1e-3 * abs(gcoef[r]) + 1e-8


def test_nucleus_attraction_to_own_cell(oracle_engine):
    # test/test_contact.jl:122-133
    J = [[1.0, 1.0], [1.0, 1.0]]
    Jrel = [[0.0, -2.0], [-2.0, 0.0]]
    terms = (F.Overlap(J, related=Jrel),)
    partial = two_body_sim(oracle_engine, terms, d=0.9, R0=(1.0, 0.4), type=(1, 2), parent=(1, 1),
                           ntypes=2, ntheta=400)
    assert partial.grad.gx[1] > 0.1
    assert partial.grad.gx[0] < -0.1
    inside = two_body_sim(oracle_engine, terms, d=0.3, R0=(1.0, 0.4), type=(1, 2), parent=(1, 1),
                          ntypes=2, ntheta=400)
    assert abs(inside.grad.gx[1]) < 1e-10
    assert abs(inside.grad.gx[0]) < 1e-10


def skin_sim(engine, skin, v0=1.0):
    # test/test_contact.jl:136-140
    b, box = F.hexagonal_monolayer(6, 6, packing=0.97, seed=2)
    return F.Simulation(b, box, types=(F.BodyType(v0=v0, Dr=1.0, Dt=0.01, muab=1.0),),
                        energies=(F.Tension(0.1), F.Overlap(1.0), F.LinearRepulsion(0.3)), dt=0.01,
                        skin=skin, seed=3, engine=engine)


def test_verlet_skin(oracle_engine):
    # test/test_contact.jl:135-161
    try:
        skin_sim(oracle_engine, -0.1)
        assert False, "negative skin accepted"
    except F.ArgumentError:
        pass
    sim = skin_sim(oracle_engine, 0.2)
    fn = overlap_fn(1.0, lin=0.3)
    for _ in range(60):
        F.step(sim)
        F.update_neighbors(sim)
        F.compute_gradients(sim)
        geo, bod = sim.geom, sim.bodies
        ref = brute_force_pressure(bod, geo.R, geo.lab, sim.box, fn, 40)
        np.testing.assert_allclose(sim.grad.P, ref, rtol=1e-12, atol=1e-12)
    assert 3 < sim.counters().builds < 50
    assert F.step(skin_sim(oracle_engine, 0.0), 10).counters().builds == 10
    a = F.step(skin_sim(oracle_engine, 0.0, v0=0.3), 100)
    b = F.step(skin_sim(oracle_engine, 0.5, v0=0.3), 100)
    assert b.counters().builds < a.counters().builds
    np.testing.assert_allclose(b.bodies.x, a.bodies.x, rtol=1e-10)
    np.testing.assert_allclose(b.bodies.coef, a.bodies.coef, rtol=1e-10, atol=1e-14)


def test_build_lists_standalone_matches_simulation(oracle_engine):
    b, box = jittered_monolayer()
    sim = F.Simulation(b, box, types=(F.BodyType(),), energies=(F.Overlap(1.0),), engine=oracle_engine)
    F.update_neighbors(sim, force=True)
    cutoff = 2 * 1.0 * (1 + math.sqrt(2 / 3)) + 0.3
    out = oracle.build_lists(b.x, b.y, sim.geom.rmax, box.Lx, box.Ly, cutoff, 0.3)
    cl, nl = sim.cells, sim.neighbors
    for k in ("cellid", "perm", "start", "stop"):
        assert np.array_equal(out[k], getattr(cl, k))
    assert np.array_equal(out["count"], nl.count)
    assert np.array_equal(out["idx"], nl.idx)
