"""Oracle geometry and body terms vs closed forms, direct evaluation and
finite differences (test/test_geometry.jl, test/test_energies.jl)."""
import math

import numpy as np
import pytest

import fcc_hip as F
import oracle
from conftest import deformed_coef
from helpers import rdirect_body, contour_direct


def test_box_helpers():
    # test/test_geometry.jl:16-21
    assert oracle.wrap(-0.5, 2.0) == (1.5, -1)
    assert oracle.wrap(4.5, 2.0) == (0.5, 2)
    assert oracle.wrap(-1e-18, 2.0) == (0.0, 0)
    assert abs(oracle.minimum_image(1.6, 2.0) + 0.4) < 1e-15
    # round is ties-to-even (src/box.jl:14): 1.0 -> 1 - 2*round(0.5) = 1 - 0
    assert oracle.minimum_image(1.0, 2.0) == 1.0
    assert oracle.minimum_image(3.0, 2.0) == -1.0


def synth_sim(engine, ntheta=40):
    # test/test_geometry.jl:23-47: one deformed, rotated body
    coef = np.zeros((11, 1))
    coef[1:6, 0] = [0.1, -0.05, 0.03, 0.0, 0.02]
    coef[6:11, 0] = [0.02, 0.04, 0.0, -0.01, 0.01]
    coef[0, 0] = math.sqrt(1 - np.sum(coef[1:, 0] ** 2) / 2)
    b = F.Bodies([0.0], [0.0], [0.7], coef, [1.3])
    return F.Simulation(b, F.PeriodicBox(20.0, 20.0), types=(F.BodyType(),), ntheta=ntheta,
                        engine=engine), coef


def test_contour_synthesis_vs_direct(oracle_engine):
    sim, coef = synth_sim(oracle_engine)
    g = sim.geom
    phi = 2 * math.pi * np.arange(40) / 40
    Rd = lambda p: rdirect_body(coef[:, 0], 0.7, 1.3, p)
    np.testing.assert_allclose(g.R[:, 0], Rd(phi), rtol=1e-13)
    h = 1e-6
    np.testing.assert_allclose(g.dR[:, 0], (Rd(phi + h) - Rd(phi - h)) / (2 * h), rtol=1e-7, atol=1e-9)
    # Julia's ≈ on arrays is norm-wise (the FD second difference is noise-limited)
    fd2 = (Rd(phi + h) - 2 * Rd(phi) + Rd(phi - h)) / h ** 2
    assert np.linalg.norm(g.ddR[:, 0] - fd2) <= 1e-3 * np.linalg.norm(fd2)
    # and exactly against the analytic second derivative -R0 Σ n² (...)
    K = 5
    dd = np.zeros(40)
    for m in range(1, K + 1):
        n = m + 1
        dd -= 1.3 * n * n * (coef[m, 0] * np.cos(n * (phi - 0.7)) + coef[K + m, 0] * np.sin(n * (phi - 0.7)))
    np.testing.assert_allclose(g.ddR[:, 0], dd, rtol=1e-12, atol=1e-13)
    # off-grid evaluation from the lab-frame coefficients (direct cos/sin)
    for psi in (-2.0, 0.3, 4.0):
        assert abs(contour_direct(g.lab[:, 0], 1.3, 5, psi) - Rd(psi)) < 1e-13
    assert g.rmax[0] >= np.max(Rd(np.linspace(0, 2 * math.pi, 1000)))


def test_basis_orthogonality_via_synthesis(oracle_engine):
    # test/test_geometry.jl:9-12 (Cᵀ C Δφ = diag(2π, π, ...)) checked on the
    # synthesised contours: ∮ R² dφ = 2π R0² (a0² + Σ(an²+bn²)/2) for any shape.
    sim, coef = synth_sim(oracle_engine)
    R = sim.geom.R[:, 0]
    lhs = np.sum(R ** 2) * 2 * math.pi / 40
    rhs = 2 * math.pi * 1.3 ** 2 * (coef[0, 0] ** 2 + np.sum(coef[1:, 0] ** 2) / 2)
    assert abs(lhs - rhs) < 1e-12
    with pytest.raises(F.ArgumentError):
        F.Simulation(F.Bodies([0.0], [0.0], [0.0], coef, [1.0]), F.PeriodicBox(5.0, 5.0),
                     types=(F.BodyType(),), ntheta=12, engine=oracle_engine)


def single_body_sim(engine, terms, theta=0.4, R0=1.2, ntheta=40, type=1, ntypes=1):
    # test/test_energies.jl:29-35
    coef = deformed_coef().reshape(-1, 1)
    b = F.Bodies([1.0], [2.0], [theta], coef, [R0], type=type)
    sim = F.Simulation(b, F.PeriodicBox(20.0, 20.0), types=tuple(F.BodyType() for _ in range(ntypes)),
                       energies=terms, ntheta=ntheta, engine=engine)
    return sim, coef


def fd_shape_gradient(sim, coef, h=1e-6):
    # test/test_energies.jl:38-53
    g = np.zeros(coef.shape[0] - 1)
    b0 = sim.bodies
    for r in range(1, coef.shape[0]):
        E = []
        for d in (h, -h):
            b = b0.copy()
            b.coef = coef.copy()
            b.coef[r, 0] += d
            sim.bodies = b
            F.refresh(sim)
            E.append(F.body_energy(sim))
        g[r - 1] = (E[0] - E[1]) / (2 * h)
    b = b0.copy()
    b.coef = coef.copy()
    sim.bodies = b
    F.refresh(sim)
    return g


def test_circle_energies(oracle_engine):
    # test/test_energies.jl:65-71
    sim, _ = single_body_sim(oracle_engine, (F.Tension(0.1),), R0=1.0, ntheta=64)
    b = sim.bodies
    b.coef[:, 0] = 0
    b.coef[0, 0] = 1
    sim.bodies = b
    F.refresh(sim)
    assert abs(F.body_energy(sim) - 0.1 * 2 * math.pi) < 1e-13
    assert abs(F.shape_index(sim)[0] - 2 * math.sqrt(math.pi)) < 1e-13


@pytest.mark.parametrize("term", [F.Tension(0.37), F.PerimeterElasticity(0.5, 3.9), F.Bending(0.2)],
                         ids=["Tension", "PerimeterElasticity", "Bending"])
def test_fd_gradient(oracle_engine, term):
    # test/test_energies.jl:73-85, rtol 1e-6; single-body terms exert no force or torque
    sim, coef = single_body_sim(oracle_engine, (term,))
    F.refresh(sim)
    g = sim.grad
    np.testing.assert_allclose(g.gcoef[1:, 0], fd_shape_gradient(sim, coef), rtol=1e-6)
    assert g.gx[0] == 0 and g.gy[0] == 0 and g.gth[0] == 0


def test_perimeter_elasticity_closed_form(oracle_engine):
    # E = λ (Q - p0 √π R0)² with Q the perimeter (body_terms.jl:28-34)
    sim, _ = single_body_sim(oracle_engine, (F.PerimeterElasticity(0.5, 3.9),))
    F.refresh(sim)
    Q = F.perimeters(sim)[0]
    assert abs(F.body_energy(sim) - 0.5 * (Q - 3.9 * math.sqrt(math.pi) * 1.2) ** 2) < 1e-13


def test_on_types_restricts_terms(oracle_engine):
    # test/test_energies.jl:96-103
    for type, active in ((1, False), (2, True)):
        sim, _ = single_body_sim(oracle_engine, (F.OnTypes(F.Tension(1.0), 2),), type=type, ntypes=2)
        F.refresh(sim)
        assert (F.body_energy(sim) > 0) == active
        assert bool(np.any(sim.grad.gcoef != 0)) == active


def test_contours_of_round_and_deformed_bodies(oracle_engine):
    # contours (src/observables.jl:36-44): a round body of radius R0 has its
    # nodes at (x + R0 cos phi_k, y + R0 sin phi_k); a deformed one at R(phi_k)
    # of the direct series around its centre
    coef = np.zeros((11, 2))
    coef[0, 0] = 1.0
    coef[1:6, 1] = [0.1, -0.05, 0.03, 0.0, 0.02]
    coef[6:11, 1] = [0.02, 0.04, 0.0, -0.01, 0.01]
    coef[0, 1] = math.sqrt(1 - np.sum(coef[1:, 1] ** 2) / 2)
    b = F.Bodies([3.0, 11.0], [4.0, 12.5], [0.0, 0.7], coef, [1.5, 1.3])
    sim = F.Simulation(b, F.PeriodicBox(20.0, 20.0), types=(F.BodyType(),), ntheta=40, engine=oracle_engine)
    X, Y = F.contours(sim)
    phi = 2 * math.pi * np.arange(40) / 40
    np.testing.assert_allclose(X[:, 0], 3.0 + 1.5 * np.cos(phi), rtol=0, atol=1e-14)
    np.testing.assert_allclose(Y[:, 0], 4.0 + 1.5 * np.sin(phi), rtol=0, atol=1e-14)
    R = np.array([rdirect_body(coef[:, 1], 0.7, 1.3, p) for p in phi])
    np.testing.assert_allclose(X[:, 1], 11.0 + R * np.cos(phi), rtol=0, atol=1e-13)
    np.testing.assert_allclose(Y[:, 1], 12.5 + R * np.sin(phi), rtol=0, atol=1e-13)
