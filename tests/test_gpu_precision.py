"""Error of the FP32 and mixed-precision (MFMA hi+lo) paths against the FP64
oracle on identical input states.

Metric (the contact density is a Heaviside of the penetration depth, so a node
whose depth is within rounding of zero may flip): nodes whose contact density
differs from the oracle's are counted as flips and reported separately; on the
bodies without a flipped node, the per-body relative error of the force
(gx, gy) and torque is |Δ| / max(|ref|, scale), where scale is the summed
magnitude of the body's node contributions Σk |Pk| Δφ |nk| (nk the outward
normal element (R cos + R' sin, R sin - R' cos), resp. |R R'| for the torque):
the natural error scale of a sum whose terms cancel (a nucleus inside its cell
feels -2 at every node and a net force ≈ 0). It must stay below 1e-4, the
force tolerance of the mixed path. Shape index and shape-mode gradients are
checked too.
"""
import math
import sys

import numpy as np
import pytest

import fcc_hip as F
from helpers import force_scales, plain_force_errors, rel_err

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def deformed_state(n=32, maxmode=8, packing=1.0, seed=11, amp=0.03):
    b, box = F.hexagonal_monolayer(n, n, packing=packing, maxmode=maxmode, seed=seed)
    rng = np.random.default_rng(seed)
    N = len(b)
    b.x = np.mod(b.x + 0.05 * rng.standard_normal(N), box.Lx)
    b.y = np.mod(b.y + 0.05 * rng.standard_normal(N), box.Ly)
    b.coef[1:, :] = amp * rng.standard_normal((b.coef.shape[0] - 1, N))
    b.R0 = np.exp(0.05 * rng.standard_normal(N))
    return b, box


def gradients(b, box, precision, engine=None, energies=None, **kw):
    energies = energies or (F.Tension(0.01), F.Overlap(1.0))
    sim = F.Simulation(b, box, types=(F.BodyType(v0=0.1, Dr=0.01),), energies=energies,
                       precision=precision, debug=True, engine=engine, **kw)
    F.refresh(sim)
    return sim, sim.grad


# two_lanes: the two-lane kernels (taken from 131072 bodies on) at this size
@pytest.mark.parametrize("two_lanes", [False, True])
@pytest.mark.parametrize("precision", ["fp32", "mixed_fp16", "mixed_bf16"])
@pytest.mark.parametrize("maxmode", [6, 8])
def test_force_error_vs_fp64(oracle_engine, precision, maxmode, two_lanes):
    say("precision", precision, "maxmode", maxmode, "two lanes" if two_lanes else "")
    b, box = deformed_state(maxmode=maxmode)
    ref_sim, ref = gradients(b, box, "fp64", engine=oracle_engine)
    dev_sim, dev = gradients(b, box, precision, two_lanes=two_lanes)
    flips = dev.P != ref.P
    nflip = int(np.count_nonzero(flips))
    nodes_in_contact = int(np.count_nonzero(ref.P))
    ok = ~np.any(flips, axis=0)
    fs, ts = force_scales(ref_sim, ref)
    ef = np.maximum(rel_err(dev.gx, ref.gx, fs), rel_err(dev.gy, ref.gy, fs))[ok]
    et = rel_err(dev.gth, ref.gth, ts)[ok]
    ec = np.max(np.abs(dev.gcoef[1:] - ref.gcoef[1:])) / np.max(np.abs(ref.gcoef[1:]))
    p_dev, p_ref = F.shape_index(dev_sim), F.shape_index(ref_sim)
    ep = np.max(np.abs(p_dev - p_ref) / p_ref)
    med, q99, pmax, field = plain_force_errors(dev, ref, ok)
    say(f"  flips {nflip} of {nodes_in_contact} contact nodes; force {ef.max():.2e}, torque {et.max():.2e}, "
        f"gcoef {ec:.2e}, shape index {ep:.2e}; bodies compared {ok.sum()} of {ok.size}")
    say(f"  plain |dF|/|F| per body: median {med:.2e}, 99 % {q99:.2e}, max {pmax:.2e}; ||dF||/||F|| {field:.2e}")
    assert field <= 1e-4
    assert nflip <= max(2, nodes_in_contact // 1000)
    assert ok.sum() >= 0.99 * ok.size
    assert ef.max() <= 1e-4
    assert et.max() <= 1e-4
    assert ec <= 1e-4
    assert ep <= 1e-5


@pytest.mark.parametrize("two_lanes", [False, True])
@pytest.mark.parametrize("precision", ["mixed_fp16", "mixed_bf16"])
def test_mixed_with_bending_and_types(oracle_engine, precision, two_lanes):
    # R'' on the matrix cores (Bending) and two coupling classes
    say("mixed bending", precision, "two lanes" if two_lanes else "")
    cells, box = F.hexagonal_monolayer(12, 12, packing=0.95, maxmode=8, seed=3)
    b = F.with_nuclei(cells, ratio=0.4)
    rng = np.random.default_rng(4)
    b.coef[1:, :] = 0.03 * rng.standard_normal((b.coef.shape[0] - 1, len(b)))
    energies = (F.OnTypes(F.Tension(0.01), 1), F.OnTypes(F.PerimeterElasticity(0.1, 3.7), 1),
                F.OnTypes(F.Bending(0.01), 2),
                F.Overlap([[1.0, 1.0], [1.0, 1.0]], related=[[0.0, -2.0], [-2.0, 0.0]]))
    types = (F.BodyType(v0=0.06, Dr=0.01), F.BodyType(Dr=0.01, drag="parent"))
    ref = F.refresh(F.Simulation(b, box, types=types, energies=energies, debug=True, engine=oracle_engine))
    dev = F.refresh(F.Simulation(b, box, types=types, energies=energies, debug=True, precision=precision,
                                 two_lanes=two_lanes))
    gr, gd = ref.grad, dev.grad
    ok = ~np.any(gd.P != gr.P, axis=0)
    assert ok.sum() >= 0.98 * ok.size
    fs, ts = force_scales(ref, gr)
    ef = np.maximum(rel_err(gd.gx, gr.gx, fs), rel_err(gd.gy, gr.gy, fs))[ok]
    ec = np.max(np.abs(gd.gcoef[1:] - gr.gcoef[1:])) / np.max(np.abs(gr.gcoef[1:]))
    eE = np.max(np.abs(gd.Ebody - gr.Ebody) / np.abs(gr.Ebody))
    say(f"  force {ef.max():.2e}, gcoef {ec:.2e}, energy {eE:.2e}")
    assert ef.max() <= 1e-4
    assert ec <= 1e-4
    assert eE <= 1e-4  # BF16 hi+lo carries ~16 significant bits (Bending weights R'' by n² ≤ 64)


@pytest.mark.parametrize("precision", ["mixed_fp16", "fp32"])
def test_mixed_graph_equals_eager(precision):
    say("mixed graph vs eager", precision)
    def mk():
        b, box = F.hexagonal_monolayer(16, 16, packing=1.0, maxmode=8, seed=5)
        return F.Simulation(b, box, types=(F.BodyType(v0=2.0, Dr=0.5),),
                            energies=(F.Tension(0.01), F.Overlap(1.0)), skin=0.05, precision=precision)
    a, c = mk(), mk()
    F.step(a, 120)
    F.run(c, 120)
    ba, bc = a.bodies, c.bodies
    assert a.counters().builds > 5
    assert np.array_equal(ba.x, bc.x) and np.array_equal(ba.coef, bc.coef)
    assert a.counters().clock == c.counters().clock == 120


@pytest.mark.parametrize("precision", ["mixed_fp16", "mixed_bf16"])
def test_force_error_at_full_size(oracle_engine, precision):
    # The headline workload itself: hexagonal_monolayer(512, 512; packing =
    # 1.0), N = 262144, modes 2..8, Tension(0.01) + Overlap(1.0), v0 0.1,
    # Dr 0.01, evolved 300 mixed steps on the device (deformed, compressed
    # contacts); the FP64 oracle and the mixed path then evaluate the
    # gradients of that same state.
    import oracle
    say("full size", precision)
    b, box = F.hexagonal_monolayer(512, 512, packing=1.0, maxmode=8, seed=1234)
    kw = dict(types=(F.BodyType(v0=0.1, Dr=0.01, muab=0.1),), energies=(F.Tension(0.01), F.Overlap(1.0)),
              dt=0.005, seed=1234, skin=0.3)
    run = F.Simulation(b, box, precision=precision, **kw)
    F.run(run, 300)
    state = run.bodies
    run.close()
    oracle.set_threads(16)
    ref_sim = F.refresh(F.Simulation(state, box, debug=True, engine=oracle_engine, **kw))
    dev_sim = F.refresh(F.Simulation(state, box, debug=True, precision=precision, **kw))
    ref, dev = ref_sim.grad, dev_sim.grad
    flips = dev.P != ref.P
    nflip, nodes_in_contact = int(np.count_nonzero(flips)), int(np.count_nonzero(ref.P))
    ok = ~np.any(flips, axis=0)
    fs, ts = force_scales(ref_sim, ref)
    ef = np.maximum(rel_err(dev.gx, ref.gx, fs), rel_err(dev.gy, ref.gy, fs))[ok]
    et = rel_err(dev.gth, ref.gth, ts)[ok]
    # a flipped node changes its body's contour gradient by the coupling J:
    # shape-mode gradients are compared on the bodies without a flip
    ec = np.max(np.abs(dev.gcoef[1:, ok] - ref.gcoef[1:, ok])) / np.max(np.abs(ref.gcoef[1:, ok]))
    med, q99, pmax, field = plain_force_errors(dev, ref, ok)
    say(f"  flips {nflip} of {nodes_in_contact} contact nodes; force {ef.max():.2e}, torque {et.max():.2e}, "
        f"gcoef {ec:.2e}; bodies compared {ok.sum()} of {ok.size}")
    say(f"  plain |dF|/|F| per body: median {med:.2e}, 99 % {q99:.2e}, max {pmax:.2e}; ||dF||/||F|| {field:.2e}")
    assert field <= 1e-4
    assert nodes_in_contact > 1_000_000
    assert nflip <= nodes_in_contact // 5000  # BF16 hi+lo: R to ~16 bits, FP16 hi+lo: ~22 bits
    assert ok.sum() >= 0.995 * ok.size
    assert ef.max() <= 1e-4
    assert et.max() <= 1e-4
    assert ec <= 1e-4


def test_polydisperse_fp32_vs_fp64(oracle_engine):
    # The plain-FP32 workload's state family at a smaller size: lognormal R0
    # (mean 1, sigma 0.1), modes 2..6 uniform and rescaled so that
    # sum sqrt(an^2 + bn^2) = 0.3, packing 0.95, Philox seed 7.
    say("polydisperse fp32")
    b, box = F.hexagonal_monolayer(64, 64, packing=0.95, maxmode=6, seed=7)
    N = len(b)
    u = F.philox_uniforms(7, 3 * N + b.coef.shape[0] * N)
    z = np.sqrt(-2 * np.log(u[:N])) * np.cos(2 * np.pi * u[N:2 * N])
    b.R0 = np.exp(0.1 * z - 0.005)
    c = 2 * u[3 * N:3 * N + (b.coef.shape[0] - 1) * N].reshape(b.coef.shape[0] - 1, N) - 1
    K = (b.coef.shape[0] - 1) // 2
    amp = np.sum(np.sqrt(c[:K] ** 2 + c[K:] ** 2), axis=0)
    b.coef[1:, :] = c * (0.3 / amp)
    ref_sim, ref = gradients(b, box, "fp64", engine=oracle_engine)
    dev_sim, dev = gradients(b, box, "fp32")
    flips = dev.P != ref.P
    ok = ~np.any(flips, axis=0)
    fs, ts = force_scales(ref_sim, ref)
    ef = np.maximum(rel_err(dev.gx, ref.gx, fs), rel_err(dev.gy, ref.gy, fs))[ok]
    et = rel_err(dev.gth, ref.gth, ts)[ok]
    ec = np.max(np.abs(dev.gcoef[1:] - ref.gcoef[1:])) / np.max(np.abs(ref.gcoef[1:]))
    say(f"  flips {int(np.count_nonzero(flips))} of {int(np.count_nonzero(ref.P))}; force {ef.max():.2e}, "
        f"torque {et.max():.2e}, gcoef {ec:.2e}")
    assert ok.sum() >= 0.99 * ok.size
    assert ef.max() <= 1e-4 and et.max() <= 1e-4 and ec <= 1e-4
