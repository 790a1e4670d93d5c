"""Contours of more than 64 nodes on the device (one workgroup per body,
k_forces_wide) vs the FP64 oracle: the gradients of test/test_cuda.jl:10-20
at nθ = 1100 (and other sizes on either side of the 64- and 2048-node
switches), the two-body contact tests of test/test_contact.jl:88-133 at their
nθ of 8000, 200 and 400, a short trajectory, graph mode == eager and
ensembles == replicas run alone, bitwise."""
import math
import sys

import numpy as np
import pytest

import fcc_hip as F
from test_gpu_parity import LAYOUT_ENERGIES, layout_bodies, same, soft_fluid
from test_oracle_contact import two_body_sim

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def layout_sim(engine, ntheta, precision="fp64", **kw):
    # test/test_layout.jl:21-29
    b, box = layout_bodies()
    types = (F.BodyType(v0=0.2, Dr=0.1), F.BodyType(Dr=0.1, drag="parent"))
    return F.Simulation(b, box, types=types, energies=LAYOUT_ENERGIES, ntheta=ntheta, precision=precision,
                        seed=7, skin=0.3, debug=True, engine=engine, **kw)


@pytest.mark.parametrize("ntheta", [65, 300, 1100, 2049])
def test_gradients_match_oracle_fp64(hip, oracle_engine, ntheta):
    # test/test_cuda.jl:10-20 (nθ = 1100 "fallback kernels"), test/test_layout.jl:88
    # (nθ = 300): rtol 1e-10, atol 1e-10
    say("wide gradients fp64 ntheta", ntheta)
    a = F.refresh(layout_sim(oracle_engine, ntheta))
    b = F.refresh(layout_sim(hip, ntheta))
    ga, gb = a.grad, b.grad
    assert np.count_nonzero(ga.P > 0) > 100
    np.testing.assert_allclose(b.geom.R, a.geom.R, rtol=1e-12, atol=1e-12)
    for f in ("P", "GR", "GD", "GDD", "gx", "gy", "gth", "gcoef", "Ebody"):
        np.testing.assert_allclose(getattr(gb, f), getattr(ga, f), rtol=1e-10, atol=1e-10, err_msg=f)


@pytest.mark.parametrize("precision", ["fp32", "mixed_fp16"])
def test_gradients_match_oracle_single(hip, oracle_engine, precision):
    # FP32 forces (the mixed mode keeps its FP64 shape constraint in the
    # update): errors relative to the largest value of each field
    say("wide gradients", precision)
    a = F.refresh(layout_sim(oracle_engine, 300))
    b = F.refresh(layout_sim(hip, 300, precision))
    ga, gb = a.grad, b.grad
    for f in ("gx", "gy", "gth", "gcoef", "Ebody"):
        ref, dev = getattr(ga, f), getattr(gb, f)
        err = np.max(np.abs(dev - ref)) / np.max(np.abs(ref))
        say(f"  {f}: max |d| / max |ref| = {err:.2e}")
        assert err < 2e-4, f


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_two_discs_exact_force(hip, oracle_engine, precision):
    # test/test_contact.jl:88-96 at nθ = 8000 (four waves per body):
    # dA/dd = -√(4 - d²)
    say("two discs nθ = 8000", precision)
    d = 1.5
    ref = two_body_sim(oracle_engine, (F.Overlap(1.0),), d=d, ntheta=8000)
    sim = F.Simulation(ref.bodies, F.PeriodicBox(10.0, 10.0), types=(F.BodyType(),),
                       energies=(F.Overlap(1.0),), ntheta=8000, precision=precision, debug=True)
    g = F.refresh(sim).grad
    exact = math.sqrt(4 - d * d)
    assert abs(g.gx[0] / exact - 1) < 1e-3
    assert abs(g.gx[1] / -exact - 1) < 1e-3
    assert abs(g.gy[0]) < (1e-6 if precision == "fp64" else 1e-4)
    if precision == "fp64":
        np.testing.assert_array_equal(g.P, ref.grad.P)
        np.testing.assert_allclose(g.gx, ref.grad.gx, rtol=1e-10)
        np.testing.assert_allclose(g.gcoef, ref.grad.gcoef, rtol=1e-10, atol=1e-10)


def test_linear_repulsion_two_bodies(hip, oracle_engine):
    # test/test_contact.jl:98-120 at nθ = 200: the oracle's gradients are
    # pinned there by finite differences of the field energy
    say("linear repulsion two bodies nθ = 200")
    ref = two_body_sim(oracle_engine, (F.LinearRepulsion(1.0),), d=1.6, deformed=True, ntheta=200)
    dev = two_body_sim(None, (F.LinearRepulsion(1.0),), d=1.6, deformed=True, ntheta=200)
    depth = ref.grad.P[(ref.grad.P != 0)]
    assert depth.size > 10
    for f in ("gx", "gy", "gth", "gcoef", "Ebody"):
        np.testing.assert_allclose(getattr(dev.grad, f), getattr(ref.grad, f), rtol=1e-10, atol=1e-12,
                                   err_msg=f)


def test_nucleus_attraction_to_own_cell(hip):
    # test/test_contact.jl:122-133 at nθ = 400
    say("nucleus attraction nθ = 400")
    J = [[1.0, 1.0], [1.0, 1.0]]
    Jrel = [[0.0, -2.0], [-2.0, 0.0]]
    terms = (F.Overlap(J, related=Jrel),)
    partial = two_body_sim(None, terms, d=0.9, R0=(1.0, 0.4), type=(1, 2), parent=(1, 1), ntypes=2,
                           ntheta=400)
    assert partial.grad.gx[1] > 0.1
    assert partial.grad.gx[0] < -0.1
    inside = two_body_sim(None, terms, d=0.3, R0=(1.0, 0.4), type=(1, 2), parent=(1, 1), ntypes=2,
                          ntheta=400)
    assert abs(inside.grad.gx[1]) < 1e-10
    assert abs(inside.grad.gx[0]) < 1e-10


def test_short_trajectory_matches_oracle_fp64(hip, oracle_engine):
    # test/test_cuda.jl:31-39 at nθ = 300: 30 steps, rtol 1e-8, same rebuild count
    say("wide trajectory fp64")
    a, b = layout_sim(oracle_engine, 300), layout_sim(hip, 300)
    F.step(a, 30)
    F.step(b, 30)
    ba, bb = a.bodies, b.bodies
    np.testing.assert_allclose(bb.x, ba.x, rtol=1e-8)
    np.testing.assert_allclose(bb.theta, ba.theta, rtol=1e-8)
    np.testing.assert_allclose(bb.coef, ba.coef, rtol=1e-8, atol=1e-12)
    assert b.counters().builds == a.counters().builds


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_graph_mode_equals_eager(prec):
    # test/test_cuda.jl:41-60 at nθ = 100: many rebuilds, frozen chunks
    say("wide graph vs eager", prec)
    a, b = soft_fluid(prec, ntheta=100), soft_fluid(prec, ntheta=100)
    F.step(a, 150)
    F.run(b, 150)
    assert a.counters().builds > 10
    assert same(a, b)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ensemble_equals_replicas_alone(precision):
    say("wide ensemble", precision)
    reps = []
    for r, (nx, ny) in enumerate([(4, 4), (6, 6), (8, 10)]):
        b, box = F.hexagonal_monolayer(nx, ny, packing=0.95, seed=20 + r)
        reps.append(F.Replica(b, box, types=(F.BodyType(v0=1.0 + r, Dr=0.5),),
                              energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005, seed=20 + r, skin=0.05))
    seq = [F.run(F.Simulation(rp.bodies, rp.box, types=rp.types, energies=rp.energies, dt=rp.dt, seed=rp.seed,
                              skin=rp.skin, precision=precision, ntheta=100), 80) for rp in reps]
    ens = F.Ensemble(reps, ntheta=100, precision=precision)
    ens.run(80)
    for a, b in zip(seq, ens.members):
        assert same(a, b)
