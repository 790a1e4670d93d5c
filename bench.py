#!/usr/bin/env python3
"""Body-steps/s of the FourierContourCells time step on one MI355X.

Workload (mixed precision at N = 262144): hexagonal_monolayer(512, 512;
packing = 1.0), N = 262144 circles, shape modes n = 2..8 (ncoef 15), nθ = 40,
BodyType(v0 = 0.1, Dr = 0.01, μab = 0.1), Tension(0.01) + Overlap(1.0),
dt = 0.005, Philox seed 1234, Verlet skin 0.3. A step is one pass of the time
step over all bodies; steps run in graph mode (captured hipGraph replayed in
chunks, one host sync per chunk) with the state resident in HBM.

Prints one JSON line: the mixed-precision path's throughput (`value`), the
plain FP32 and the mixed BF16 paths' and the same workload with a forces and
an update kernel per step for comparison, the roofline of the dominant kernel
of the timed run (the one-kernel step: its FP32 vector work against the
vector peak, its contractions against the matrix-core peak, its compulsory
bytes against HBM) from HIP events over the timed run (fcc_run_timed), the
forces part as its own kernel (back-to-back at the end state,
fcc_time_kernels) and the update kernel of a two-kernel step, the list-build
stages, the FP64 CPU baseline (the oracle, OpenMP over bodies) on a bounded
sample of the same workload, and the batched-ensemble workload (64 replicas
of uneven size in one launch per kernel and step) next to the same replicas
run as independent simulations, and the two N = 65536 workloads (`workloads`:
a polydisperse FP32 monolayer, cells with nuclei in mixed precision and plain
FP32) over the same steps. The timed run and the kernel timing are
roctx ranges ("timed", "kernel_timing") for rocprofv3 --marker-trace. Fields
read from committed profiles name their file (`traffic_source`,
`issue_source`).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fouriercontourcells.jl_amd"))

import ctypes  # noqa: E402

import numpy as np  # noqa: E402

import fcc_hip as F  # noqa: E402

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 (vendor figure; the vector rate assumes packed FP32)
PEAK_FP32_UNPACKED_TFLOPS = 78.65  # the vector rate without packed FP32 (this build compiles without it)
PEAK_FP16_MFMA_TFLOPS = 2500.0     # MI355X dense FP16 / BF16 matrix cores
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
DTYPES = {"fp64": "f64", "fp32": "f32", "mixed": "mixed_fp16", "mixed_fp16": "mixed_fp16",
          "mixed_bf16": "mixed_bf16"}
TRAFFIC_FILE = "profiles/timestep_N262144_hbm_traffic.txt"
ISSUE_FILE = "profiles/forces_N262144_sq_counters.txt"


def marker(name=None):
    """Open (name) or close (None) a roctx range through torch, so that a
    rocprofv3 --marker-trace of this run delimits the timed region."""
    try:
        import torch
        if name is None:
            torch.cuda.nvtx.range_pop()
        else:
            torch.cuda.nvtx.range_push(name)
    except Exception:
        pass


def make_sim(n, precision, seed=1234, two_kernels=False):
    b, box = F.hexagonal_monolayer(n, n, packing=1.0, maxmode=8, seed=seed)
    return F.Simulation(b, box, types=(F.BodyType(v0=0.1, Dr=0.01, muab=0.1),),
                        energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005, ntheta=40,
                        precision=precision, seed=seed, skin=0.3, two_kernels=two_kernels)


def polydisperse_sim(precision, n=256, seed=7, two_kernels=False):
    """Polydisperse monolayer: hexagonal_monolayer(256, 256; packing 0.95),
    N = 65536, lognormal R0 (σ_log 0.1, mean 1), modes 2..6 with aₙ, bₙ
    uniform in [-1, 1] rescaled to Σ√(aₙ² + bₙ²) = 0.3 (a₀ from the area
    constraint at construction), v0 = 0.06, Dr = 0.01, μab = 0.1,
    Tension(0.01) + Overlap(1.0), dt = 0.005, seed 7 (numpy's generator with
    the same seed for R0 and the shapes)."""
    b, box = F.hexagonal_monolayer(n, n, packing=0.95, maxmode=6, seed=seed)
    rng = np.random.default_rng(seed)
    N = len(b)
    sig = 0.1
    b.R0 = np.exp(sig * rng.standard_normal(N) - sig * sig / 2)
    K = (b.coef.shape[0] - 1) // 2
    ab = rng.uniform(-1.0, 1.0, size=(2 * K, N))
    amp = np.sqrt(ab[:K] ** 2 + ab[K:] ** 2).sum(axis=0)
    b.coef[1:, :] = ab * (0.3 / amp)
    return F.Simulation(b, box, types=(F.BodyType(v0=0.06, Dr=0.01, muab=0.1),),
                        energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005, ntheta=40,
                        precision=precision, seed=seed, skin=0.3, two_kernels=two_kernels)


def nuclei_sim(precision, nx=256, ny=128, seed=1, two_kernels=False):
    """Typed bodies with nuclei: hexagonal_monolayer(256, 128; packing 0.95)
    cells + one nucleus each (with_nuclei, ratio 0.4), 65536 bodies, modes
    2..8; cells: Tension(0.01), PerimeterElasticity(0.1, 3.7), v0 = 0.06;
    nuclei: Bending(0.01), drag = parent; Overlap(J = 1, related = [0 -2; -2 0])
    on all; Dr = 0.01, dt = 0.002."""
    cells, box = F.hexagonal_monolayer(nx, ny, packing=0.95, maxmode=8, seed=seed)
    b = F.with_nuclei(cells, ratio=0.4)
    types = (F.BodyType(v0=0.06, Dr=0.01, muab=0.1), F.BodyType(Dr=0.01, muab=0.1, drag="parent"))
    energies = (F.OnTypes(F.Tension(0.01), 1), F.OnTypes(F.PerimeterElasticity(0.1, 3.7), 1),
                F.OnTypes(F.Bending(0.01), 2),
                F.Overlap([[1.0, 1.0], [1.0, 1.0]], related=[[0.0, -2.0], [-2.0, 0.0]]))
    return F.Simulation(b, box, types=types, energies=energies, dt=0.002, ntheta=40,
                        precision=precision, seed=seed, skin=0.3, two_kernels=two_kernels)


WORKLOADS = {
    "polydisperse65536": (polydisperse_sim, "fp32"),
    "nuclei65536": (nuclei_sim, "mixed"),
}


def bench_workload(name, precision, steps, warmup):
    """Throughput of one of the other workloads (its own JSON line), with the
    plain FP32 path beside a mixed one, and the workload's precision with two
    kernels per step ("<precision>_two_kernels"; the same as the first leg
    where a step is two kernels anyway, e.g. with parent drag)."""
    maker, default = WORKLOADS[name]
    precision = precision or default
    out = {"workload": name, "doc": " ".join(maker.__doc__.split()), "precision": precision}
    legs = [(precision, False)] + ([("fp32", False)] if precision != "fp32" else []) + [(precision, True)]
    for prec, tk in legs:
        sim = maker(prec, two_kernels=tk)
        if tk:
            prec = prec + "_two_kernels"
        F.run(sim, warmup)
        clock_warmup(sim)
        sec, _, _ = timed_run(sim, steps)
        fms, ums = kernel_durations(sim)
        v = sim.N * steps / sec
        print(f"[bench] {name} {prec}: {v / 1e6:.1f} M body-steps/s ({sec / steps * 1e3:.4f} ms/step; "
              f"forces {fms * 1e3:.1f} us, update {ums * 1e3:.1f} us), N={sim.N}, "
              f"builds {sim.counters().builds}", file=sys.stderr, flush=True)
        out[prec] = {"value": v, "unit": "body-steps/s", "N": sim.N, "steps": steps,
                     "ms_per_step": sec / steps * 1e3, "forces_ms": fms, "update_ms": ums}
        sim.close()
    return out


def cpu_small(steps=1000):
    """FP64 oracle on the small CPU workload: hexagonal_monolayer(32, 32;
    packing 0.95), N = 1024, modes 2..6, BodyType(v0 = 0.06, Dr = 0.01,
    μab = 0.1), Tension(0.01) + Overlap(1.0), dt = 0.005, θ from Philox seed
    1234, noise seed 1, `steps` steps on all host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    cores = max(1, min(cores, len(os.sched_getaffinity(0))))
    out = {}
    for c in sorted({1, cores}):
        oracle.set_threads(c)
        b, box = F.hexagonal_monolayer(32, 32, packing=0.95, maxmode=6, seed=1234)
        sim = F.Simulation(b, box, types=(F.BodyType(v0=0.06, Dr=0.01, muab=0.1),),
                           energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005, ntheta=40, seed=1,
                           skin=0.3, engine=oracle.engine())
        t0 = time.perf_counter()
        F.step(sim, steps)
        dt = time.perf_counter() - t0
        out[f"threads_{c}"] = {"value": sim.N * steps / dt, "ms_per_step": dt / steps * 1e3}
        print(f"[bench] CPU oracle N=1024, {c} threads: {sim.N * steps / dt / 1e6:.3f} M body-steps/s "
              f"({dt / steps * 1e3:.3f} ms/step)", file=sys.stderr, flush=True)
    return out


def sync(sim):
    import torch
    torch.cuda.synchronize()
    sim.counters()  # synchronises the simulation's own stream


def timed_run(sim, steps):
    """Graph-mode run of `steps` steps, timed on the host between two device
    synchronisations (torch.cuda.synchronize waits for every stream of the
    device, the simulation's own included); HIP events around each chunk's
    graph launches (one-kernel steps) or around the kernels of each chunk's
    eagerly launched last step (two-kernel steps) give the average kernel
    durations over the same region (fcc_run_timed)."""
    import torch
    f, u = ctypes.c_double(), ctypes.c_double()
    sync(sim)
    t0 = time.perf_counter()
    rc = sim.engine.run_timed(sim._h, int(steps), ctypes.byref(f), ctypes.byref(u))
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError(sim.engine.last_error())
    return sec, f.value, u.value


def kernel_durations(sim, reps=50):
    """Average forces / update kernel durations (ms) from HIP events on the
    simulation's stream around `reps` back-to-back forces launches and `reps`
    back-to-back steps (fcc_time_kernels; the state is restored afterwards)."""
    f, u = ctypes.c_double(), ctypes.c_double()
    rc = sim.engine.time_kernels(sim._h, int(reps), ctypes.byref(f), ctypes.byref(u))
    if rc != 0:
        raise RuntimeError(sim.engine.last_error())
    return f.value, u.value


def clock_warmup(sim):
    """State-preserving kernel work (fcc_time_kernels restores the state it
    advances) before a timed region: an idle GPU runs at lower clocks for the
    first milliseconds of work, longer than a few warm-up steps take."""
    kernel_durations(sim)


def binning_durations(sim, reps=20):
    """Cell-list and neighbour-list build of the current state, `reps` times
    (fcc_time_rebuild): average milliseconds of the stages cell index, scan,
    scatter, cell sort, neighbours, from HIP events recorded by the dispatches
    of each stage's first and last kernel (kernel durations)."""
    out = (ctypes.c_double * 5)()
    rc = sim.engine.time_rebuild(sim._h, int(reps), out)
    if rc != 0:
        raise RuntimeError(sim.engine.last_error())
    return list(out)


def binning_roofline(sim, stage_ms, builds, steps):
    """Algorithmic HBM bytes per stage of one list build (N bodies, C cells,
    S = total neighbour entries, r = bytes of a bound: 4 in FP32 / mixed, 8 in
    FP64) and the achieved GB/s: cell index 4C + 28N (count clear; x, y
    read, cell id write, count atomic), scan 20C (two passes), scatter 16N,
    cell sort 8C + 8N + 2(16 + r)N (segments sorted in place; x, y, rmax
    gathered into cell-sorted copies), neighbours 8C + (16 + r)N own state +
    (16 + r)N sorted copies read once + 4(S + N) perm of the candidates in
    range + 4S list + 4N counts + (16 + r)N Verlet reference (with a skin)."""
    N = sim.N
    cells = sim.cells
    C = cells.ncx * cells.ncy
    S = int(sim.neighbors.count.sum())
    r = 8 if sim.precision == "fp64" else 4
    names = ["cell_index", "scan", "scatter", "cell_sort", "neighbors"]
    nbytes = [4 * C + 28 * N, 20 * C, 16 * N, 8 * C + 8 * N + 2 * (16 + r) * N,
              8 * C + 3 * (16 + r) * N + 4 * (S + N) + 4 * S + 4 * N]
    stages = {}
    for n, ms, b in zip(names, stage_ms, nbytes):
        stages[n] = {"ms": ms, "bytes": b, "achieved": b / (ms * 1e-3) / 1e9, "unit": "GB/s",
                     "frac": b / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS}
    total = sum(stage_ms)
    return {"bound": "hbm", "peak": PEAK_HBM_GBS, "cells": C, "list_entries": S, "stages": stages,
            "ms_per_build": total, "builds_per_step": builds / steps,
            "ms_per_step_amortised": total * builds / steps}


def calibrate_copy(nbytes=1 << 30, reps=5):
    """A plain device copy of known size (torch), to scale the FETCH_SIZE /
    WRITE_SIZE counters of a rocprofv3 --pmc run of this same mode."""
    import torch
    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": nbytes, "reps": reps}), flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (TRAFFIC_FILE), or None."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == kernel:
            return int(f[1]) + int(f[2])
    return None


def measured_issue(kernel):
    """(VALU issue fraction, MFMA busy fraction) of `kernel` from the committed
    rocprofv3 --pmc SQ counter summary (ISSUE_FILE), or (None, None)."""
    path = os.path.join(ROOT, ISSUE_FILE)
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == kernel:
            return float(f[1]), float(f[2])
    return None, None


def contact_pairs(sim):
    """(tested, in-range) (node, neighbour) pairs of the current state: the
    reference tests every listed neighbour at every node and evaluates R_j(ψ)
    for those within rmax_j (src/energies/contact.jl:67-79)."""
    geo, bod, nl = sim.geom, sim.bodies, sim.neighbors
    nth, N = sim.ntheta, sim.N
    phi = 2 * math.pi * np.arange(nth) / nth
    ux = bod.x[None, :] + geo.R * np.cos(phi)[:, None]
    uy = bod.y[None, :] + geo.R * np.sin(phi)[:, None]
    Lx, Ly = sim.box.Lx, sim.box.Ly
    tested = int(nl.count.sum()) * nth
    inr = 0
    for q in range(int(nl.count.max())):
        sel = np.flatnonzero(nl.count > q)
        j = nl.idx[q, sel] - 1
        dx = ux[:, sel] - bod.x[j][None, :]
        dx -= Lx * np.round(dx / Lx)
        dy = uy[:, sel] - bod.y[j][None, :]
        dy -= Ly * np.round(dy / Ly)
        inr += int(np.count_nonzero(dx * dx + dy * dy <= (geo.rmax[j] ** 2)[None, :]))
    return tested, inr


# Compulsory HBM bytes per body of a one-kernel step (k_step_mfma /
# k_step_fast2, FP32 state): reads x, y 16, θ 4, lab 60, R0 4, type 4, rmax,
# rmin 8, Verlet reference x0, y0, rmax0 20, neighbour count and 6 list
# entries 28; writes the new x, y 16, θ 4, lab 60, rmax, rmin 8 and the
# gradients glab 60, gx, gy, gθ 12, Ebody 4. Neighbour gathers are L2 reads.
STEP_BYTES = 148 + 164
# The update kernel of a two-kernel step: 188 B read + 92 B written per body.
UPDATE_BYTES = 280


def forces_flops(sim, tested, inr):
    """Algorithmic FLOPs of one forces launch as (contractions, vector work):
    contour contractions 2·[2·(3nθ)·ncoef] per body (synthesis + pullback,
    src/geometry.jl:113-122, src/gradients.jl:374-386; on the matrix cores in
    the mixed modes), and the vector work: contact 12 per tested and 7 + 8K
    per in-range (node, neighbour) pair, 10 per node for the Tension
    integrand and its gradient."""
    N, nth, nc, K = sim.N, sim.ntheta, sim.ncoef, (sim.ncoef - 1) // 2
    return N * 2 * (2 * 3 * nth * nc), 12 * tested + (7 + 8 * K) * inr + 10 * nth * N


def ensemble_replicas(nrep=64, seed=4):
    """Batched-ensemble workload: `nrep` replicas with N log-uniform in
    [1024, 16384] (nx ≈ ny ≈ √N, ny even), packing U[0.9, 1.0], v0 linear in
    0.02..0.2, Dr cycling over {0.005, 0.01, 0.02}, Tension(η ~ U[0.005, 0.05])
    + Overlap(1.0), modes 2..6, nθ = 40, dt = 0.005, Philox key = replica id
    (sizes and η from numpy's generator with `seed`)."""
    rng = np.random.default_rng(seed)
    reps = []
    for r in range(nrep):
        n = math.exp(rng.uniform(math.log(1024), math.log(16384)))
        nx = max(2, round(math.sqrt(n)))
        ny = max(2, 2 * round(n / nx / 2))
        packing = rng.uniform(0.9, 1.0)
        eta = rng.uniform(0.005, 0.05)
        v0 = 0.02 + (0.2 - 0.02) * r / max(1, nrep - 1)
        Dr = (0.005, 0.01, 0.02)[r % 3]
        b, box = F.hexagonal_monolayer(nx, ny, packing=packing, maxmode=6, seed=r)
        reps.append(F.Replica(b, box, types=(F.BodyType(v0=v0, Dr=Dr, muab=0.1),),
                              energies=(F.Tension(eta), F.Overlap(1.0)), dt=0.005, seed=r, skin=0.3))
    return reps


def bench_ensemble(steps, warmup, independent_steps, workers=16, independent=True):
    """Ensemble throughput (all replicas in one launch per kernel and step,
    graph mode, launch shape from the autotuner) next to the same replicas
    as independent simulations, each run by its own host thread on its own
    stream (the reference's one-task-per-simulation pattern)."""
    import torch
    from concurrent.futures import ThreadPoolExecutor
    reps = ensemble_replicas()
    total = sum(len(rp.bodies) for rp in reps)
    ens = F.Ensemble(reps, precision="fp32")
    ens.run(warmup)
    launch, tune_ms = ens.autotune(20)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run(steps)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    value = total * steps / sec
    builds = sum(m.counters().builds for m in ens.members)
    # kernel breakdown from a short timed run after the measured one (the
    # timed form runs each chunk's last step eagerly between events)
    fms, ums = ens.run_timed(max(1, steps // 10))
    ens.close()
    print(f"[bench] ensemble: {len(reps)} replicas, {total} bodies, {value / 1e6:.1f} M body-steps/s "
          f"({sec / steps * 1e3:.4f} ms/step; forces {fms * 1e3:.1f} us, update {ums * 1e3:.1f} us), "
          f"launch {launch}, {builds} rebuilds", file=sys.stderr, flush=True)
    out = {"value": value, "unit": "body-steps/s", "replicas": len(reps), "bodies": total,
           "steps": steps, "ms_per_step": sec / steps * 1e3, "forces_ms": fms, "update_ms": ums,
           "launch": launch, "autotune_ms_per_step": tune_ms, "rebuilds": builds,
           "workload": "64 replicas, N log-uniform [1024,16384], packing U[0.9,1], v0 0.02..0.2, "
                       "Dr {0.005,0.01,0.02}, Tension(U[0.005,0.05])+Overlap(1), modes 2..6, FP32, "
                       "dt=0.005, key = replica id"}
    if not independent:
        return out
    sims = [F.Simulation(rp.bodies, rp.box, types=rp.types, energies=rp.energies, dt=rp.dt, ntheta=40,
                         precision="fp32", seed=rp.seed, skin=rp.skin) for rp in reps]
    with ThreadPoolExecutor(max_workers=workers) as pool:
        list(pool.map(lambda s: F.run(s, warmup), sims))
        for s in sims:
            s.counters()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        list(pool.map(lambda s: F.run(s, independent_steps), sims))
        for s in sims:
            s.counters()
        torch.cuda.synchronize()
        isec = time.perf_counter() - t0
    ivalue = total * independent_steps / isec
    for s in sims:
        s.close()
    print(f"[bench] independent simulations ({workers} host threads, one stream each): "
          f"{ivalue / 1e6:.1f} M body-steps/s", file=sys.stderr, flush=True)
    out.update({"independent_value": ivalue, "independent_steps": independent_steps,
                "independent_threads": workers, "speedup_vs_independent": value / ivalue})
    return out


def cpu_baseline(n, seconds=15.0):
    """FP64 oracle (C restatement of the reference CPU path, OpenMP over
    bodies) on the same workload: 1 warm-up step, then steps until about
    `seconds` have elapsed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    cores = max(1, min(cores, len(os.sched_getaffinity(0))))
    oracle.set_threads(cores)
    b, box = F.hexagonal_monolayer(n, n, packing=1.0, maxmode=8, seed=1234)
    sim = F.Simulation(b, box, types=(F.BodyType(v0=0.1, Dr=0.01, muab=0.1),),
                       energies=(F.Tension(0.01), F.Overlap(1.0)), dt=0.005, ntheta=40, seed=1234,
                       skin=0.3, engine=oracle.engine())
    F.step(sim, 1)
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        F.step(sim, 1)
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": sim.N * steps / dt, "unit": "body-steps/s", "cores": cores, "kind": "port",
            "sample": f"N={sim.N}, FP64 oracle, {steps} steps after 1 warm-up step, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=512, help="bodies per lattice row (N = n²)")
    ap.add_argument("--precision", default="mixed")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=2000)
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only the ensemble leg (its own JSON line)")
    ap.add_argument("--no-two-kernels", action="store_true",
                    help="without the leg that runs the workload with two kernels per step")
    ap.add_argument("--no-workloads", action="store_true",
                    help="without the polydisperse and nuclei workloads (N=65536) of the full line")
    ap.add_argument("--no-independent", action="store_true",
                    help="ensemble leg without the independent-simulations comparison")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", choices=sorted(WORKLOADS),
                    help="measure another workload instead (its own JSON line)")
    ap.add_argument("--cpu-small", action="store_true",
                    help="only the FP64 oracle on the N=1024 CPU workload (its own JSON line)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the time step has no multi-GPU path "
              "(one simulation runs on one device); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.cpu_small:
        print(json.dumps({"cpu_small": cpu_small()}), flush=True)
        return
    if args.workload:
        prec = None if args.precision == "mixed" else args.precision
        print(json.dumps(bench_workload(args.workload, prec, args.steps, args.warmup)), flush=True)
        return

    if args.ensemble_only:
        print(json.dumps(bench_ensemble(args.ensemble_steps, args.warmup, max(1, args.ensemble_steps // 4),
                                        independent=not args.no_independent)), flush=True)
        return

    print(f"[bench] workload N={args.n ** 2} precision={args.precision}", file=sys.stderr, flush=True)
    sim = make_sim(args.n, args.precision)
    F.run(sim, args.warmup)
    clock_warmup(sim)
    tested0, inr0 = contact_pairs(sim)
    clock_warmup(sim)
    marker("timed")
    sec, run_fms, run_ums = timed_run(sim, args.steps)
    marker()
    # kernel timing right after the timed run (an idle GPU clocks down; the
    # pair count below is host work), at its end state
    marker("kernel_timing")
    end_fms, end_ums = kernel_durations(sim)
    marker()
    tested1, inr1 = contact_pairs(sim)
    tested, inr = (tested0 + tested1) / 2, (inr0 + inr1) / 2
    c = sim.counters()
    binning = binning_roofline(sim, binning_durations(sim), c.builds, args.warmup + args.steps)
    print("[bench] list build " + ", ".join(f"{k} {v['ms'] * 1e3:.1f} us" for k, v in binning["stages"].items())
          + f" = {binning['ms_per_build'] * 1e3:.1f} us per build", file=sys.stderr, flush=True)
    value = sim.N * args.steps / sec
    print(f"[bench] {args.precision}: {value / 1e6:.1f} M body-steps/s ({sec / args.steps * 1e3:.4f} ms/step), "
          f"builds {c.builds}", file=sys.stderr, flush=True)

    def other_path(prec):
        s2 = make_sim(args.n, prec)
        F.run(s2, args.warmup)
        clock_warmup(s2)
        psec, _, _ = timed_run(s2, args.steps)
        v = s2.N * args.steps / psec
        print(f"[bench] {prec}: {v / 1e6:.1f} M body-steps/s", file=sys.stderr, flush=True)
        s2.close()
        return v

    plain = None if args.no_plain or args.precision == "fp32" else other_path("fp32")
    bf16 = other_path("mixed_bf16") if args.precision in ("mixed", "mixed_fp16") and not args.no_plain else None

    # the same workload with a forces and an update kernel per step
    # (FCC_FLAG_TWO_KERNELS): its throughput, and the update kernel's own
    # duration (eager samples over its timed run)
    two = None
    if not args.no_two_kernels:
        s3 = make_sim(args.n, args.precision, two_kernels=True)
        F.run(s3, args.warmup)
        clock_warmup(s3)
        tsec, tfms, tums = timed_run(s3, args.steps)
        two = {"value": s3.N * args.steps / tsec, "ms_per_step": tsec / args.steps * 1e3,
               "forces_ms": tfms, "update_ms": tums}
        print(f"[bench] {args.precision}, two kernels per step: {two['value'] / 1e6:.1f} M body-steps/s "
              f"(forces {tfms * 1e3:.1f} us, update {tums * 1e3:.1f} us)", file=sys.stderr, flush=True)
        s3.close()

    mixed = args.precision.startswith("mixed")
    # kernel durations sampled over the timed run: flops of its mean state
    fl_mfma, fl_valu = forces_flops(sim, tested, inr)
    if not mixed:  # the contractions run on the vector units too
        fl_valu, fl_mfma = fl_valu + fl_mfma, 0.0
    one = run_ums == 0.0  # the timed run's steps were one kernel each
    fkernel = "k_forces_mfma" if mixed else "k_forces_fast2"
    skernel = ("k_step_mfma" if mixed else "k_step_fast2") if one else fkernel
    t_f = run_fms * 1e-3
    a_valu, a_mfma = fl_valu / t_f / 1e12, fl_mfma / t_f / 1e12
    on_profile = sim.N == 262144
    valu_frac, mfma_frac = measured_issue(skernel) if on_profile else (None, None)
    fvalu_frac, fmfma_frac = measured_issue(fkernel) if on_profile else (None, None)
    step_bytes = STEP_BYTES * sim.N
    upd_bytes = UPDATE_BYTES * sim.N
    upd_ms = two["update_ms"] if two else (run_ums if not one else None)
    print(f"[bench] step kernel {run_fms * 1e3:.1f} us (samples over the timed run); forces kernel alone, "
          f"back-to-back at its end state: {end_fms * 1e3:.1f} us; pairs tested {tested}, in range {inr}",
          file=sys.stderr, flush=True)

    ensemble = None
    if not args.no_ensemble:
        ensemble = bench_ensemble(args.ensemble_steps, min(args.warmup, 200), max(1, args.ensemble_steps // 4),
                                  independent=not args.no_independent)

    # the other workloads, each with its paths, over the same steps
    workloads = None
    if not args.no_workloads:
        workloads = {name: bench_workload(name, None, args.steps, args.warmup) for name in sorted(WORKLOADS)}

    cpu = None if args.no_cpu else cpu_baseline(args.n, args.cpu_seconds)
    timing_note = ("HIP events on the simulation's stream around the graph launches of each chunk of the "
                   "timed run, the chunk's time over its steps (one-kernel steps; two-kernel steps: around the "
                   "kernels of each chunk's last step, launched eagerly with the graph's kernels and arguments), "
                   "fcc_run_timed, roctx range 'timed'; flops of the mean of its first and last state")
    end_note = ("captured graph of 50 back-to-back launches at the end state of the timed run, right after it, "
                "with ~400 untimed launches first (fcc_time_kernels; roctx range 'kernel_timing'); the forces "
                "series re-evaluates one state, whose data stay in L2 between launches")
    two_note = ("HIP events around the eager last step of each chunk of the same workload run with a forces and "
                "an update kernel per step (FCC_FLAG_TWO_KERNELS, fcc_run_timed)")
    fa_valu, fa_mfma = fl_valu / (end_fms * 1e-3) / 1e12, fl_mfma / (end_fms * 1e-3) / 1e12

    line = {
        "metric": "body-steps/sec at N=262144 (mixed vs plain FP32, 1 GPU) @ force err ≤1e-4",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES.get(args.precision, args.precision),
        "data": "synthetic",
        "config": {"workload": f"hexagonal_monolayer({args.n},{args.n}; packing=1.0), N={sim.N}, "
                               f"modes 2..8, ntheta=40, Tension(0.01)+Overlap(1.0), v0=0.1, Dr=0.01, "
                               f"dt=0.005, skin=0.3, graph mode",
                   "precision": args.precision, "N": sim.N},
        "plain_fp32_value": plain,
        "mixed_over_plain": (value / plain) if plain else None,
        "mixed_bf16_value": bf16,
        "two_kernels_value": two["value"] if two else None,
        "one_over_two_kernels": (value / two["value"]) if two else None,
        "roofline": {"kernel": skernel,
                     "what": ("one kernel per step: forces and update of 64 bodies per workgroup" if one
                              else "forces kernel of a two-kernel step"),
                     "bound": "valu", "achieved": a_valu,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": a_valu / PEAK_FP32_TFLOPS,
                     "peak_unpacked": PEAK_FP32_UNPACKED_TFLOPS,
                     "frac_unpacked": a_valu / PEAK_FP32_UNPACKED_TFLOPS,
                     "traffic": measured_traffic(skernel) if on_profile else None,
                     "traffic_source": TRAFFIC_FILE if on_profile else None,
                     "launch_ms": run_fms, "launch_ms_source": timing_note,
                     "pairs_tested": tested, "pairs_in_range": inr,
                     "flops_per_launch": fl_valu + fl_mfma,
                     "parts": {
                         "valu": {"what": "contact + body terms (FP32 vector)", "flops": fl_valu,
                                  "achieved": a_valu, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                  "frac": a_valu / PEAK_FP32_TFLOPS},
                         "mfma": ({"what": "contour synthesis + pullback (FP16/BF16 hi+lo matrix cores; "
                                           "algorithmic flops counted once)",
                                   "flops": fl_mfma, "achieved": a_mfma, "peak": PEAK_FP16_MFMA_TFLOPS,
                                   "unit": "TFLOP/s", "frac": a_mfma / PEAK_FP16_MFMA_TFLOPS}
                                  if mixed else None),
                         "hbm": ({"what": f"compulsory state and gradient bytes of the step ({STEP_BYTES} B "
                                          "per body)", "bytes": step_bytes,
                                  "achieved": step_bytes / t_f / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                  "frac": step_bytes / t_f / 1e9 / PEAK_HBM_GBS} if one else None)},
                     "valu_issue_frac": valu_frac, "mfma_busy_frac": mfma_frac,
                     "issue_source": ISSUE_FILE if on_profile else None},
        "roofline_forces": {"kernel": fkernel,
                            "what": "the forces part as its own kernel (fcc_compute_gradients, ensembles, "
                                    "two-kernel steps), same flops",
                            "bound": "valu", "achieved": fa_valu, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                            "frac": fa_valu / PEAK_FP32_TFLOPS,
                            "frac_unpacked": fa_valu / PEAK_FP32_UNPACKED_TFLOPS,
                            "mfma": ({"achieved": fa_mfma, "peak": PEAK_FP16_MFMA_TFLOPS,
                                      "frac": fa_mfma / PEAK_FP16_MFMA_TFLOPS} if mixed else None),
                            "traffic": measured_traffic(fkernel) if on_profile else None,
                            "traffic_source": TRAFFIC_FILE if on_profile else None,
                            "launch_ms": end_fms, "launch_ms_source": end_note,
                            "valu_issue_frac": fvalu_frac, "mfma_busy_frac": fmfma_frac,
                            "issue_source": ISSUE_FILE if on_profile else None},
        "roofline_update": ({"kernel": "k_update", "bound": "hbm",
                             "what": "update kernel of a two-kernel step (ensembles, parent drag, bonds, "
                                     "FCC_FLAG_TWO_KERNELS)",
                             "traffic": measured_traffic("k_update") if on_profile else None,
                             "traffic_source": TRAFFIC_FILE if on_profile else None,
                             "achieved": upd_bytes / (upd_ms * 1e-3) / 1e9, "peak": PEAK_HBM_GBS,
                             "unit": "GB/s", "frac": upd_bytes / (upd_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                             "bytes_per_launch": upd_bytes, "launch_ms": upd_ms,
                             "launch_ms_source": two_note if two else timing_note}
                            if upd_ms else None),
        "roofline_binning": binning,
        "cpu_baseline": cpu,
        "ensemble": ensemble,
        "workloads": workloads,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
