"""Host-side mirror of the FourierContourCells.jl API for the time step.

Names, argument meaning and errors follow the reference (Julia `!` functions
become plain names: `step!` -> `step`, `run!` -> `run`,
`compute_gradients!` -> `compute_gradients`, `update_neighbors!` ->
`update_neighbors`). The time step itself runs in libfcc_hip.so through the C
ABI of include/fcc.h; this module only validates, converts and copies.

Reference: src/simulation.jl, src/bodies.jl, src/box.jl, src/initial.jl,
src/observables.jl, src/energies/*.jl.
"""
import ctypes
import math
from dataclasses import dataclass, field
from types import SimpleNamespace

import numpy as np

from . import _abi
from ._abi import (FCC_MAX_TERMS, FCC_MAX_TYPES, FccBodies, FccConfig)


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


class UnsupportedError(RuntimeError):
    """A configuration the reference accepts but that has no native kernel
    (the Julia extension falls back to the reference path for it)."""


class DeviceError(RuntimeError):
    """HIP runtime error."""


_ERRORS = {
    _abi.FCC_ERR_ARGUMENT: ArgumentError,
    _abi.FCC_ERR_DIMENSION: DimensionMismatch,
    _abi.FCC_ERR_UNSUPPORTED: UnsupportedError,
    _abi.FCC_ERR_DEVICE: DeviceError,
    _abi.FCC_ERR_STATE: RuntimeError,
}


def _check(engine, rc):
    if rc != 0:
        raise _ERRORS.get(rc, RuntimeError)(engine.last_error())


# ------------------------------------------------------------------ types


@dataclass(frozen=True)
class BodyType:
    """BodyType (src/bodies.jl:30-45)."""
    v0: float = 0.0
    Dr: float = 0.0
    Dt: float = 0.0
    mur: float = 1.0
    mutheta: float = 1.0
    muab: float = 0.1
    drag: str = "substrate"

    def __post_init__(self):
        if self.drag not in ("substrate", "parent"):
            raise ArgumentError(f"drag must be :substrate or :parent, got :{self.drag}")


@dataclass(frozen=True)
class PeriodicBox:
    """PeriodicBox (src/box.jl:6-11)."""
    Lx: float
    Ly: float


class EnergyTerm:
    kind = 0
    types = ()

    def params(self):
        return (0.0, 0.0)


@dataclass(frozen=True)
class Tension(EnergyTerm):
    """Line tension η·l (src/energies/body_terms.jl:8-18)."""
    eta: float
    kind = _abi.FCC_TERM_TENSION

    def params(self):
        return (self.eta, 0.0)


@dataclass(frozen=True)
class PerimeterElasticity(EnergyTerm):
    """λ (l - p0 √π R0)² (src/energies/body_terms.jl:28-34)."""
    lam: float
    p0: float
    kind = _abi.FCC_TERM_PERIMETER_ELASTICITY

    def params(self):
        return (self.lam, self.p0)


@dataclass(frozen=True)
class Bending(EnergyTerm):
    """κ ∮ (R² + 2R'² - R R'')² / (R² + R'²)^{5/2} dφ (test/test_energies.jl:6-10)."""
    kappa: float
    kind = _abi.FCC_TERM_BENDING

    def params(self):
        return (self.kappa, 0.0)


@dataclass(frozen=True)
class Overlap(EnergyTerm):
    """Overlap(J; related = J) (src/energies/contact.jl:12-28)."""
    J: object
    related: object = None
    kind = _abi.FCC_TERM_OVERLAP

    def tables(self):
        J = np.atleast_2d(np.asarray(self.J, dtype=np.float64))
        R = J if self.related is None else np.atleast_2d(np.asarray(self.related, dtype=np.float64))
        if J.shape != R.shape or J.shape[0] != J.shape[1]:
            raise DimensionMismatch("coupling tables must be square and equal-sized")
        return J, R


@dataclass(frozen=True)
class LinearRepulsion(EnergyTerm):
    """g(d) = k·d for d > 0 (src/energies/contact.jl:37-41)."""
    k: float
    kind = _abi.FCC_TERM_LINEAR_REPULSION

    def params(self):
        return (self.k, 0.0)


@dataclass(frozen=True)
class HarmonicBond(EnergyTerm):
    """HarmonicBond(k; r0 = 0): (k/2)(|r_i - r_j| - r0)^2 between bonded
    centres (src/energies/bonds.jl:12-29); acts on the `bonds` pairs."""
    k: float
    r0: float = 0.0
    kind = _abi.FCC_TERM_HARMONIC_BOND

    def params(self):
        return (self.k, self.r0)


def parent_bonds(bodies):
    """parent_bonds (src/energies/bonds.jl:37): one bond (i, parent) per body
    that is not its own parent, 1-based."""
    return [(i + 1, int(p)) for i, p in enumerate(bodies.parent) if p != i + 1]


@dataclass(frozen=True)
class OnTypes(EnergyTerm):
    """OnTypes(term, types...) (src/energies/interface.jl:278-284)."""
    term: EnergyTerm
    on: tuple = field(default_factory=tuple)

    def __init__(self, term, *types):
        object.__setattr__(self, "term", term)
        object.__setattr__(self, "on", tuple(int(t) for t in types))

    @property
    def kind(self):
        return self.term.kind


def _unwrap(term):
    on = ()
    while isinstance(term, OnTypes):
        on = term.on if not on else tuple(t for t in on if t in term.on)
        term = term.term
    return term, on


# ------------------------------------------------------------------ bodies


class Bodies:
    """Structure-of-arrays body state (src/bodies.jl:62-86), host side."""

    def __init__(self, x, y, theta, coef, R0, type=1, parent=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        N = x.shape[0]
        y = np.ascontiguousarray(y, dtype=np.float64)
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        R0 = np.ascontiguousarray(R0, dtype=np.float64)
        coef = np.asfortranarray(coef, dtype=np.float64)
        if not (y.shape[0] == theta.shape[0] == R0.shape[0] == N) or coef.ndim != 2 or coef.shape[1] != N:
            raise DimensionMismatch("inconsistent body array lengths")
        self.x, self.y, self.theta, self.R0, self.coef = x, y, theta, R0, coef
        if np.isscalar(type):
            self.type = np.full(N, int(type), dtype=np.int32)
        else:
            self.type = np.ascontiguousarray(type, dtype=np.int32)
        self.parent = (np.arange(1, N + 1, dtype=np.int32) if parent is None
                       else np.ascontiguousarray(parent, dtype=np.int32))
        self.imgx = np.zeros(N, dtype=np.int32)
        self.imgy = np.zeros(N, dtype=np.int32)

    def __len__(self):
        return self.x.shape[0]

    def copy(self):
        b = Bodies(self.x.copy(), self.y.copy(), self.theta.copy(), self.coef.copy(),
                   self.R0.copy(), self.type.copy(), self.parent.copy())
        b.imgx[:] = self.imgx
        b.imgy[:] = self.imgy
        return b

    def struct(self):
        """fcc_bodies view (keeps the arrays alive through self)."""
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        self.coef = np.asfortranarray(self.coef)
        return FccBodies(dp(self.x), dp(self.y), dp(self.theta), ip(self.imgx), ip(self.imgy),
                         dp(self.coef), dp(self.R0), ip(self.type), ip(self.parent))


def vcat(*bs):
    """vcat(bodies...) (src/bodies.jl:96-102)."""
    offs = np.cumsum([0] + [len(b) for b in bs[:-1]])
    b = Bodies(np.concatenate([q.x for q in bs]), np.concatenate([q.y for q in bs]),
               np.concatenate([q.theta for q in bs]), np.hstack([q.coef for q in bs]),
               np.concatenate([q.R0 for q in bs]), np.concatenate([q.type for q in bs]),
               np.concatenate([q.parent + o for q, o in zip(bs, offs)]))
    b.imgx[:] = np.concatenate([q.imgx for q in bs])
    b.imgy[:] = np.concatenate([q.imgy for q in bs])
    return b


def philox_uniforms(seed, n, stream=0):
    """n uniforms in (0, 1) from Philox4x32-10 with key = seed and counter
    (i, stream, 0, 0), i = 1..n (src/random.jl:25-44, FP64 form). Used for
    reproducible initial orientations in place of Julia's Xoshiro stream."""
    M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
    c0 = np.arange(1, n + 1, dtype=np.uint64)
    c1 = np.full(n, stream & 0xFFFFFFFF, dtype=np.uint64)
    c2 = np.zeros(n, dtype=np.uint64)
    c3 = np.zeros(n, dtype=np.uint64)
    k0, k1 = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = np.uint64(M0) * c0
        p1 = np.uint64(M1) * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & mask
        hi1, lo1 = p1 >> np.uint64(32), p1 & mask
        c0, c1, c2, c3 = hi1 ^ c1 ^ np.uint64(k0), lo1, hi0 ^ c3 ^ np.uint64(k1), lo0
        k0 = (k0 + W0) & 0xFFFFFFFF
        k1 = (k1 + W1) & 0xFFFFFFFF
    j = ((c0 << np.uint64(32)) | c1) >> np.uint64(12)
    return (j.astype(np.float64) + 0.5) * 2.0 ** -52


def hexagonal_monolayer(nx, ny, packing=0.95, R0=1.0, maxmode=6, seed=1):
    """hexagonal_monolayer (src/initial.jl:11-29), orientations from
    philox_uniforms(seed) instead of Julia's Xoshiro(1)."""
    if ny % 2 != 0:
        raise ArgumentError("ny must be even for a periodic hexagonal lattice")
    N = nx * ny
    r = R0 * math.sqrt(math.pi / (2 * math.sqrt(3) * packing))
    box = PeriodicBox(2 * r * nx, math.sqrt(3) * r * ny)
    n = np.arange(N)
    row, col = np.divmod(n, nx)
    x = (2 * col + (row % 2)) * r
    y = math.sqrt(3) * r * row
    theta = 2 * math.pi * philox_uniforms(seed, N)
    coef = np.zeros((2 * maxmode - 1, N))
    coef[0, :] = 1.0
    return Bodies(x, y, theta, coef, np.full(N, float(R0))), box


def with_nuclei(cells, ratio=0.5, type=2, seed=2):
    """with_nuclei (src/initial.jl:39-53)."""
    N = len(cells)
    coef = np.zeros_like(cells.coef)
    coef[0, :] = 1.0
    nuclei = Bodies(cells.x.copy(), cells.y.copy(), 2 * math.pi * philox_uniforms(seed, N), coef,
                    ratio * cells.R0, type=type)
    nuclei.imgx[:] = cells.imgx
    nuclei.imgy[:] = cells.imgy
    b = vcat(cells, nuclei)
    b.parent[N:] = np.arange(1, N + 1, dtype=np.int32)
    return b


# -------------------------------------------------------------- simulation


def _config(bodies, box, types, energies, dt, ntheta, precision, seed, skin, flags):
    types = tuple(types)
    energies = tuple(energies)
    if len(types) > FCC_MAX_TYPES:
        raise UnsupportedError(f"at most {FCC_MAX_TYPES} body types are native")
    if len(energies) > FCC_MAX_TERMS:
        raise UnsupportedError(f"at most {FCC_MAX_TERMS} energy terms are native")
    nc = bodies.coef.shape[0]
    if nc % 2 != 1:
        raise DimensionMismatch("coefficient rows must be odd: [a₀; aₙ; bₙ]")
    cfg = FccConfig()
    cfg.n_bodies = len(bodies)
    cfg.n_modes = nc // 2
    cfg.n_theta = int(ntheta)
    cfg.n_types = len(types)
    cfg.n_terms = len(energies)
    for q, t in enumerate(types):
        if not isinstance(t, BodyType):
            raise ArgumentError("types must be BodyTypes")
        ct = cfg.types[q]
        ct.v0, ct.Dr, ct.Dt, ct.mur, ct.mutheta, ct.muab = t.v0, t.Dr, t.Dt, t.mur, t.mutheta, t.muab
        ct.parent_drag = 1 if t.drag == "parent" else 0
    for q, e in enumerate(energies):
        if not isinstance(e, EnergyTerm):
            raise ArgumentError("energy terms must be BodyTerms, ContactTerms or BondTerms")
        base, on = _unwrap(e)
        ct = cfg.terms[q]
        ct.kind = base.kind
        ct.n_on = len(on)
        for k, t in enumerate(on[:FCC_MAX_TYPES]):
            ct.on_types[k] = t
        if len(on) > FCC_MAX_TYPES:
            raise UnsupportedError("OnTypes with more than FCC_MAX_TYPES types")
        if isinstance(base, Overlap):
            J, R = base.tables()
            nt = J.shape[0]
            for a in range(FCC_MAX_TYPES):
                for b in range(FCC_MAX_TYPES):
                    # scalar couplings apply to every type pair
                    src = (a, b) if nt > 1 else (0, 0)
                    if nt == 1 or (a < nt and b < nt):
                        ct.J[a * FCC_MAX_TYPES + b] = J[src]
                        ct.Jrel[a * FCC_MAX_TYPES + b] = R[src]
        else:
            ct.p[0], ct.p[1] = base.params()
    cfg.dt = float(dt)
    cfg.Lx = float(box.Lx)
    cfg.Ly = float(box.Ly)
    cfg.skin = float(skin)
    cfg.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    cfg.precision = _abi.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    cfg.flags = int(flags)
    return cfg


class Simulation:
    """Simulation(bodies, box; types, energies, dt, nθ, T, seed, skin)
    (src/simulation.jl:66-105) on the device library.

    `precision` replaces `T`: "fp64" (T = Float64), "fp32" (T = Float32),
    "mixed" / "mixed_fp16" / "mixed_bf16" (MFMA contour contractions).
    `engine` is the C-ABI implementation (default: libfcc_hip.so).
    `two_lanes` launches the two-lane forces kernels at every size (the
    library picks four lanes per body below 131072 bodies), for tests.
    `two_kernels` keeps a step as a forces and an update kernel where the
    library would run it as one kernel (bitwise the same results), for tests
    and kernel profiles.
    """

    def __init__(self, bodies, box, types, energies=(), dt=0.005, ntheta=40, precision="fp64",
                 seed=1, skin=0.3, debug=False, engine=None, bonds=(), two_lanes=False, two_kernels=False):
        self._setup(engine if engine is not None else _abi.hip_engine(), bodies, box, types, energies,
                    dt, ntheta, precision, seed, skin, debug, two_lanes, two_kernels)
        h = ctypes.c_void_p()
        _check(self.engine, self.engine.create(ctypes.byref(self._cfg),
                                               ctypes.byref(self._bodies.struct()), ctypes.byref(h)))
        self._h = h
        if len(bonds):
            self.set_bonds(bonds)

    def checkpoint(self):
        """The run state as bytes (fcc_checkpoint_save): body state, Verlet
        reference, neighbour list, device clock and counters."""
        need = ctypes.c_int64()
        _check(self.engine, self.engine.checkpoint_save(self._h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        _check(self.engine, self.engine.checkpoint_save(self._h, buf, need.value, ctypes.byref(need)))
        return buf.raw

    def restore(self, data):
        """Continue from `checkpoint()` bytes of a simulation with the same
        configuration (fcc_checkpoint_load)."""
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        _check(self.engine, self.engine.checkpoint_load(self._h, buf, len(data)))
        return self

    def set_bonds(self, bonds):
        """The `bonds` pairs (i, j), 1-based (src/simulation.jl:14-15)."""
        pairs = np.ascontiguousarray(np.asarray(list(bonds), dtype=np.int32).reshape(-1, 2))
        _check(self.engine, self.engine.set_bonds(self._h, len(pairs),
                                                  pairs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        self.bonds = [tuple(int(v) for v in b) for b in pairs]

    def _setup(self, engine, bodies, box, types, energies, dt, ntheta, precision, seed, skin, debug,
               two_lanes=False, two_kernels=False):
        self.engine = engine
        self.box = box
        self.types = tuple(types)
        self.energies = tuple(energies)
        self.N = len(bodies)
        self.ncoef = bodies.coef.shape[0]
        self.ntheta = int(ntheta)
        self.precision = precision
        self._bodies = bodies.copy()
        self._owner = None
        flags = ((_abi.FCC_FLAG_DEBUG_BUFFERS if debug else 0) | (_abi.FCC_FLAG_TWO_LANES if two_lanes else 0)
                 | (_abi.FCC_FLAG_TWO_KERNELS if two_kernels else 0))
        self._cfg = _config(bodies, box, types, energies, dt, ntheta, precision, seed, skin, flags)

    def close(self):
        if getattr(self, "_owner", None) is not None:  # a replica: the ensemble frees it
            self._h = None
            return
        if getattr(self, "_h", None):
            self.engine.destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- state ------------------------------------------------------------
    @property
    def bodies(self):
        b = self._bodies
        _check(self.engine, self.engine.get_bodies(self._h, ctypes.byref(b.struct())))
        return b.copy()

    @bodies.setter
    def bodies(self, b):
        if len(b) != self.N or b.coef.shape[0] != self.ncoef:
            raise DimensionMismatch("body count or coefficient rows changed")
        self._bodies = b.copy()
        _check(self.engine, self.engine.set_bodies(self._h, ctypes.byref(self._bodies.struct())))

    @property
    def dt(self):
        return self._cfg.dt

    @dt.setter
    def dt(self, v):
        self._cfg.dt = float(v)
        _check(self.engine, self.engine.set_dt(self._h, float(v)))

    def counters(self):
        st, t, cl, bu = ctypes.c_int64(), ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
        _check(self.engine, self.engine.get_counters(self._h, ctypes.byref(st), ctypes.byref(t),
                                                     ctypes.byref(cl), ctypes.byref(bu)))
        return SimpleNamespace(step=st.value, t=t.value, clock=cl.value, builds=bu.value)

    @property
    def t(self):
        return self.counters().t

    @property
    def step_count(self):
        return self.counters().step

    @property
    def grad(self):
        N, nc, nth = self.N, self.ncoef, self.ntheta
        gx, gy, gth, Eb = (np.zeros(N) for _ in range(4))
        glab, gcoef = np.zeros((nc, N), order="F"), np.zeros((nc, N), order="F")
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        _check(self.engine, self.engine.get_gradients(self._h, dp(gx), dp(gy), dp(gth), dp(glab),
                                                      dp(gcoef), dp(Eb)))
        g = SimpleNamespace(gx=gx, gy=gy, gth=gth, glab=glab, gcoef=gcoef, Ebody=Eb)
        P, GR, GD, GDD = (np.zeros((nth, N), order="F") for _ in range(4))
        rc = self.engine.get_contour_gradients(self._h, dp(P), dp(GR), dp(GD), dp(GDD))
        if rc == 0:
            g.P, g.GR, g.GD, g.GDD = P, GR, GD, GDD
        return g

    @property
    def geom(self):
        N, nc, nth = self.N, self.ncoef, self.ntheta
        lab = np.zeros((nc, N), order="F")
        rmax = np.zeros(N)
        R, dR, ddR = (np.zeros((nth, N), order="F") for _ in range(3))
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        _check(self.engine, self.engine.get_geometry(self._h, dp(lab), dp(rmax), dp(R), dp(dR), dp(ddR)))
        return SimpleNamespace(lab=lab, rmax=rmax, R=R, dR=dR, ddR=ddR)

    @property
    def neighbors(self):
        cap = ctypes.c_int32()
        count = np.zeros(self.N, dtype=np.int32)
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        _check(self.engine, self.engine.get_neighbors(self._h, ctypes.byref(cap), ip(count), None))
        idx = np.zeros((cap.value, self.N), dtype=np.int32, order="F")
        _check(self.engine, self.engine.get_neighbors(self._h, ctypes.byref(cap), ip(count), ip(idx)))
        return SimpleNamespace(capacity=cap.value, count=count, idx=idx,
                               builds=self.counters().builds)

    def set_neighbor_capacity(self, cap):
        _check(self.engine, self.engine.set_neighbor_capacity(self._h, int(cap)))

    @property
    def cells(self):
        ncx, ncy = ctypes.c_int32(), ctypes.c_int32()
        _check(self.engine, self.engine.get_celllist(self._h, ctypes.byref(ncx), ctypes.byref(ncy),
                                                     None, None, None, None))
        nce = ncx.value * ncy.value
        cellid, perm = np.zeros(self.N, np.int32), np.zeros(self.N, np.int32)
        start, stop = np.zeros(nce, np.int32), np.zeros(nce, np.int32)
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        _check(self.engine, self.engine.get_celllist(self._h, ctypes.byref(ncx), ctypes.byref(ncy),
                                                     ip(cellid), ip(perm), ip(start), ip(stop)))
        return SimpleNamespace(ncx=ncx.value, ncy=ncy.value, cellid=cellid, perm=perm,
                               start=start, stop=stop)


class Replica:
    """The arguments of Simulation(...) that an ensemble replica keeps as
    its own (bodies, box, types, energies, dt, seed, skin)."""

    def __init__(self, bodies, box, types, energies=(), dt=0.005, seed=1, skin=0.3, bonds=()):
        self.bodies, self.box, self.types, self.energies = bodies, box, tuple(types), tuple(energies)
        self.dt, self.seed, self.skin, self.bonds = dt, seed, skin, list(bonds)


class Ensemble:
    """Independent simulations advanced together: the reference runs one
    Simulation per task and stream (README.md:249-261); here every step of
    all replicas is one launch per kernel, captured in a hipGraph
    (fcc_ensemble_* of include/fcc.h). Replicas share `precision`, `ntheta`
    and the shape-mode count; sizes, boxes, parameters and seeds differ.
    `members` are Simulation handles of the replicas (every function of
    this module applies to them); each replica ends bitwise where `run`
    would take it alone. `tune` runs the launch autotuner once.
    `two_lanes` and `two_kernels` are the Simulation flags, for every replica.
    """

    def __init__(self, replicas, ntheta=40, precision="fp32", debug=False, tune=False, tune_steps=20,
                 engine=None, two_lanes=False, two_kernels=False):
        self.engine = engine if engine is not None else _abi.hip_engine()
        replicas = list(replicas)
        if not replicas:
            raise ArgumentError("an ensemble needs at least one replica")
        n = len(replicas)
        self.members = []
        cfgs = (FccConfig * n)()
        structs = (FccBodies * n)()
        for r, rp in enumerate(replicas):
            sim = Simulation.__new__(Simulation)
            sim._setup(self.engine, rp.bodies, rp.box, rp.types, rp.energies, rp.dt, ntheta, precision,
                       rp.seed, rp.skin, debug, two_lanes, two_kernels)
            cfgs[r] = sim._cfg
            structs[r] = sim._bodies.struct()
            self.members.append(sim)
        h = ctypes.c_void_p()
        _check(self.engine, self.engine.ensemble_create(n, cfgs, structs, ctypes.byref(h)))
        self._h = h
        for r, sim in enumerate(self.members):
            mh = ctypes.c_void_p()
            _check(self.engine, self.engine.ensemble_member(self._h, r, ctypes.byref(mh)))
            sim._h = mh
            sim._owner = self
            if len(replicas[r].bonds):
                sim.set_bonds(replicas[r].bonds)
        if tune:
            self.autotune(tune_steps)

    def __len__(self):
        return len(self.members)

    def close(self):
        if getattr(self, "_h", None):
            for sim in self.members:
                sim._h = None
            self.engine.ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps):
        """run!(sim, nsteps) on every replica."""
        _check(self.engine, self.engine.ensemble_run(self._h, int(nsteps)))
        return self

    def run_timed(self, nsteps):
        """`run` with HIP-event timings of the chunks' last steps: (forces ms,
        update ms) per ensemble step."""
        f, u = ctypes.c_double(), ctypes.c_double()
        _check(self.engine, self.engine.ensemble_run_timed(self._h, int(nsteps), ctypes.byref(f),
                                                           ctypes.byref(u)))
        return f.value, u.value

    @property
    def launch(self):
        lc = _abi.FccLaunch()
        _check(self.engine, self.engine.ensemble_get_launch(self._h, ctypes.byref(lc)))
        return {"update_block": lc.update_block, "xcd_order": lc.xcd_order, "chunk": lc.chunk,
                "lanes": lc.lanes, "kernels": lc.kernels}

    @launch.setter
    def launch(self, v):
        lc = _abi.FccLaunch(int(v.get("update_block", 256)), int(v.get("xcd_order", 1)),
                            int(v.get("chunk", 0)), int(v.get("lanes", 0)), int(v.get("kernels", 0)))
        _check(self.engine, self.engine.ensemble_set_launch(self._h, ctypes.byref(lc)))

    def observer_start(self, nbins=256, prange=(3.5, 5.5), max_samples=1024):
        """`observer_start` on every replica."""
        c = _observer_config(nbins, prange, max_samples)
        _check(self.engine, self.engine.ensemble_observer_start(self._h, ctypes.byref(c)))
        for sim in self.members:
            sim._obs = c
        return self

    def observe(self):
        """One device sample of every replica."""
        _check(self.engine, self.engine.ensemble_observe(self._h))
        return self

    def run_observed(self, nsteps, every):
        """`run` in runs of `every` steps, every replica sampled after each."""
        _check(self.engine, self.engine.ensemble_run_observed(self._h, int(nsteps), int(every)))
        return self

    def observations(self):
        """`observations` of every replica."""
        return [observations(sim) for sim in self.members]

    def autotune(self, steps=20):
        """Time every candidate launch shape over `steps` steps from the same
        state (restored afterwards), keep the fastest; returns (launch, ms per step)."""
        lc, ms = _abi.FccLaunch(), ctypes.c_double()
        _check(self.engine, self.engine.ensemble_autotune(self._h, int(steps), ctypes.byref(lc),
                                                          ctypes.byref(ms)))
        return self.launch, ms.value


# ---------------------------------------------------------------- functions


def step(sim, n=1):
    """step!(sim, n) (src/simulation.jl:165-173)."""
    _check(sim.engine, sim.engine.step(sim._h, int(n)))
    return sim


def run(sim, nsteps, every=0, callback=None, graph=True):
    """run!(sim, nsteps; every, callback, graph) (src/simulation.jl:220-235);
    graph mode replays a captured hipGraph (ext/FourierContourCellsCUDAExt.jl:144-180)."""
    def advance(n):
        if graph and hasattr(sim.engine, "run"):
            _check(sim.engine, sim.engine.run(sim._h, int(n)))
        else:
            step(sim, n)
    if callback is None or every <= 0:
        advance(nsteps)
    else:
        done = 0
        while done < nsteps:
            n = min(every, nsteps - done)
            advance(n)
            done += n
            callback(sim)
    return sim


def compute_gradients(sim, frame=True):
    """compute_gradients!(sim) (src/simulation.jl:150-158)."""
    _check(sim.engine, sim.engine.compute_gradients(sim._h, 1 if frame else 0))
    return sim.grad


def update_neighbors(sim, force=False):
    """update_neighbors!(sim; force) (src/simulation.jl:132-142)."""
    _check(sim.engine, sim.engine.update_neighbors(sim._h, 1 if force else 0))
    return sim


def refresh(sim):
    """The reference tests' refresh! (test/runtests.jl:9-15): re-enforce the
    shape, recompute geometry, rebuild neighbours, compute gradients."""
    sim.bodies = sim.bodies
    update_neighbors(sim, force=True)
    compute_gradients(sim)
    return sim


def _observables(sim):
    N = sim.N
    per, area, p = np.zeros(N), np.zeros(N), np.zeros(N)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    _check(sim.engine, sim.engine.shape_index(sim._h, dp(per), dp(area), dp(p)))
    return per, area, p


def _select(v, sim, type):
    return v if type is None else v[sim._bodies.type == type]


def perimeters(sim, type=None):
    """perimeters (src/observables.jl:9-10)."""
    return _select(_observables(sim)[0], sim, type)


def areas(sim, type=None):
    """areas (src/observables.jl:13-14)."""
    return _select(_observables(sim)[1], sim, type)


def shape_index(sim, type=None):
    """shape_index (src/observables.jl:18)."""
    return _select(_observables(sim)[2], sim, type)


def positions(sim):
    """positions (src/observables.jl:21): 2 x N."""
    b = sim.bodies
    return np.vstack([b.x, b.y])


def unwrapped_positions(sim):
    """unwrapped_positions (src/observables.jl:24-29): 2 x N."""
    b = sim.bodies
    return np.vstack([b.x + b.imgx * sim.box.Lx, b.y + b.imgy * sim.box.Ly])


def contours(sim):
    """contours(sim) (src/observables.jl:36-44): host nθ x N matrices (X, Y)
    of contour node coordinates around the wrapped centres."""
    X, Y = (np.zeros((sim.ntheta, sim.N), order="F") for _ in range(2))
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    _check(sim.engine, sim.engine.contours(sim._h, dp(X), dp(Y)))
    return X, Y


def frame_request(sim):
    """Enqueue the contours of the current state and their copy to the host
    (fcc_frame_request); steps issued afterwards overlap the copy."""
    _check(sim.engine, sim.engine.frame_request(sim._h))
    return sim


def frame_read(sim):
    """The frame of the last `frame_request`: (X, Y, device clock)."""
    X, Y = (np.zeros((sim.ntheta, sim.N), order="F") for _ in range(2))
    clock = ctypes.c_int64()
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    _check(sim.engine, sim.engine.frame_read(sim._h, dp(X), dp(Y), ctypes.byref(clock)))
    return X, Y, clock.value


def body_energy(sim):
    """body_energy (src/observables.jl:47)."""
    return float(np.sum(sim.grad.Ebody))


# ------------------------------------------------- device observations


def _observer_config(nbins, prange, max_samples):
    pmin, pmax = (float(v) for v in prange)
    return _abi.FccObserverConfig(int(nbins), int(max_samples), pmin, pmax)


def observer_start(sim, nbins=256, prange=(3.5, 5.5), max_samples=1024):
    """Start a device-side observation of `sim` (fcc_observer_start): the
    current unwrapped positions (src/observables.jl:24-29) become the origin
    of the mean squared displacement; each later `observe` adds one sample of
    the per-type mean squared displacement, mean shape index and mean squared
    shape index (src/observables.jl:18) and bins every body's shape index
    into a per-type histogram over `prange` with `nbins` bins (plus one bin
    below and one at or above the range)."""
    c = _observer_config(nbins, prange, max_samples)
    _check(sim.engine, sim.engine.observer_start(sim._h, ctypes.byref(c)))
    sim._obs = c
    return sim


def observe(sim):
    """One device sample of the current state (no host transfer)."""
    _check(sim.engine, sim.engine.observe(sim._h))
    return sim


def run_observed(sim, nsteps, every):
    """run!(sim, nsteps; every, callback) (src/simulation.jl:220-235) with a
    device sample as the callback."""
    _check(sim.engine, sim.engine.run_observed(sim._h, int(nsteps), int(every)))
    return sim


def observations(sim):
    """The samples so far: `clock` [samples] (device step clock), `msd`,
    `p_mean`, `p2_mean` [samples x types], `hist` [types x (nbins + 2)] and
    the bin `edges` [nbins + 1] of bins 1..nbins."""
    c = getattr(sim, "_obs", None)
    if c is None:
        raise RuntimeError("no observation started (observer_start)")
    nt, ms, nb = len(sim.types), c.max_samples, c.nbins
    ns = ctypes.c_int32()
    clock = np.zeros(ms, np.int64)
    msd, pm, p2 = (np.zeros((ms, nt)) for _ in range(3))
    hist = np.zeros((nt, nb + 2), np.uint64)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    _check(sim.engine, sim.engine.observer_read(
        sim._h, ctypes.byref(ns), clock.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), dp(msd), dp(pm), dp(p2),
        hist.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
    n = ns.value
    edges = c.pmin + (c.pmax - c.pmin) * np.arange(nb + 1) / nb
    return SimpleNamespace(clock=clock[:n], msd=msd[:n], p_mean=pm[:n], p2_mean=p2[:n], hist=hist,
                           edges=edges, nbins=nb, prange=(c.pmin, c.pmax))


def philox_noise(n, seed, draw, precision="fp64", engine=None):
    """philox_noise!(zeros(T, 3, n), philox_key(seed), draw) (src/random.jl:126-130)."""
    engine = engine if engine is not None else _abi.hip_engine()
    prec = _abi.FCC_FP64 if precision == "fp64" else _abi.FCC_FP32
    out = np.zeros((3, n), dtype=np.float64 if prec == _abi.FCC_FP64 else np.float32, order="F")
    _check(engine, engine.philox_noise(int(seed), int(draw), int(n), prec,
                                       out.ctypes.data_as(ctypes.c_void_p)))
    return out
