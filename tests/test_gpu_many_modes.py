"""Shape modes beyond 16 (K = maxmode - 1 from 16 to 31: the 31-mode
instantiations of every kernel, basis rows of 64 slots) vs the FP64 oracle:
gradients at nθ = 40 and 64 (one lane per body) and above 64 (one workgroup
per body), a short trajectory, single precision, and an ensemble == its
replicas run alone, bitwise."""
import sys

import numpy as np
import pytest

import fcc_hip as F
from test_gpu_parity import LAYOUT_ENERGIES, same

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def deformed_sim(engine, maxmode, ntheta, precision="fp64", amplitude=0.02, energies=LAYOUT_ENERGIES, **kw):
    # the nucleated, deformed monolayer of test/test_layout.jl:21-29 with more modes
    cells, box = F.hexagonal_monolayer(5, 4, packing=0.97, seed=2, maxmode=maxmode)
    b = F.with_nuclei(cells)
    rng = np.random.default_rng(5)
    b.coef[1:, :] = amplitude * rng.standard_normal((b.coef.shape[0] - 1, len(b)))
    types = (F.BodyType(v0=0.2, Dr=0.1), F.BodyType(Dr=0.1, drag="parent"))
    return F.Simulation(b, box, types=types, energies=energies, ntheta=ntheta, precision=precision,
                        seed=7, skin=0.3, debug=True, engine=engine, **kw)


@pytest.mark.parametrize("maxmode,ntheta", [(17, 40), (17, 64), (24, 100), (32, 100), (32, 2100)])
def test_gradients_match_oracle_fp64(hip, oracle_engine, maxmode, ntheta):
    # test/test_cuda.jl:10-20 tolerances: rtol 1e-10, atol 1e-10
    say("many modes fp64: maxmode", maxmode, "ntheta", ntheta)
    a = F.refresh(deformed_sim(oracle_engine, maxmode, ntheta))
    b = F.refresh(deformed_sim(hip, maxmode, ntheta))
    ga, gb = a.grad, b.grad
    assert np.count_nonzero(ga.P > 0) > 100
    np.testing.assert_allclose(b.geom.R, a.geom.R, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(b.geom.ddR, a.geom.ddR, rtol=1e-11, atol=1e-11)
    for f in ("P", "GR", "GD", "GDD", "gx", "gy", "gth", "gcoef", "Ebody"):
        np.testing.assert_allclose(getattr(gb, f), getattr(ga, f), rtol=1e-10, atol=1e-10, err_msg=f)


@pytest.mark.parametrize("maxmode,ntheta,bending,dt", [(17, 40, True, 0.001), (32, 100, False, 0.005)])
def test_short_trajectory_matches_oracle_fp64(hip, oracle_engine, maxmode, ntheta, bending, dt):
    # test/test_cuda.jl:31-39: 30 steps, rtol 1e-8, same rebuild count, on
    # states the oracle itself integrates stably: it turns a 1e-13 change of
    # the coefficients into 1e-1 within 20 steps where the explicit update of
    # the stiff high modes is unstable (deformations of 0.02 per coefficient;
    # Bending on the R0 = 0.4 nuclei at dt = 0.005 from 17 modes, at
    # dt = 0.0005 with 32), and keeps it at 1e-14 on these
    say("many modes trajectory fp64: maxmode", maxmode, "ntheta", ntheta, "bending", bending, "dt", dt)
    energies = LAYOUT_ENERGIES if bending else tuple(e for e in LAYOUT_ENERGIES
                                                     if not isinstance(getattr(e, "term", e), F.Bending))
    assert len(energies) == (4 if bending else 3)
    a = deformed_sim(oracle_engine, maxmode, ntheta, amplitude=0.01, energies=energies, dt=dt)
    b = deformed_sim(hip, maxmode, ntheta, amplitude=0.01, energies=energies, dt=dt)
    F.step(a, 30)
    F.step(b, 30)
    np.testing.assert_allclose(b.bodies.x, a.bodies.x, rtol=1e-8)
    np.testing.assert_allclose(b.bodies.theta, a.bodies.theta, rtol=1e-8)
    np.testing.assert_allclose(b.bodies.coef, a.bodies.coef, rtol=1e-8, atol=1e-12)
    assert b.counters().builds == a.counters().builds


@pytest.mark.parametrize("precision", ["fp32", "mixed_fp16"])
def test_gradients_match_oracle_single(hip, oracle_engine, precision):
    say("many modes", precision)
    a = F.refresh(deformed_sim(oracle_engine, 20, 48))
    b = F.refresh(deformed_sim(hip, 20, 48, precision))
    for f in ("gx", "gy", "gth", "gcoef", "Ebody"):
        ref, dev = getattr(a.grad, f), getattr(b.grad, f)
        err = np.max(np.abs(dev - ref)) / np.max(np.abs(ref))
        say(f"  {f}: max |d| / max |ref| = {err:.2e}")
        assert err < 2e-4, f


@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed_fp16"])
def test_ensemble_equals_replicas_alone(precision):
    say("many modes ensemble", precision)
    reps = []
    for r, (nx, ny) in enumerate([(4, 4), (6, 6), (8, 10)]):
        b, box = F.hexagonal_monolayer(nx, ny, packing=0.95, seed=30 + r, maxmode=18)
        rng = np.random.default_rng(r)
        b.coef[1:, :] = 0.01 * rng.standard_normal((b.coef.shape[0] - 1, len(b)))
        reps.append(F.Replica(b, box, types=(F.BodyType(v0=1.0 + r, Dr=0.5),),
                              energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005, seed=30 + r, skin=0.05))
    seq = [F.run(F.Simulation(rp.bodies, rp.box, types=rp.types, energies=rp.energies, dt=rp.dt, seed=rp.seed,
                              skin=rp.skin, precision=precision, ntheta=40), 80) for rp in reps]
    ens = F.Ensemble(reps, ntheta=40, precision=precision)
    ens.run(80)
    for a, b in zip(seq, ens.members):
        assert same(a, b)
