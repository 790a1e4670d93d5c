# FourierContourCellsHIPExt — package extension of FourierContourCells.jl that
# runs the time step on an AMD MI355X through libfcc_hip.so (C ABI of
# include/fcc.h). It is registered like FourierContourCellsCUDAExt, through
# [weakdeps]/[extensions] of the package's Project.toml
# (Project-additions.toml next to this folder); its weak dependency is the
# FCCHIP package (../FCCHIP), which provides `HIPBackend`, `Bending` and the
# library path.
#
#   using FourierContourCells, FCCHIP
#   sim = Simulation(bodies, box, HIPBackend(); types, energies, dt, T = Float32)
#   run!(sim, 10_000)
#   p = shape_index(sim)
#
# What runs on the reference path instead (the reference's own
# `Simulation(bodies, box; ...)` on its CPU backend, with every argument
# passed on): user-defined terms (no native kernel), `hook!` (Julia code
# between the kernels of every step, src/simulation.jl:181-191), an explicit
# `layout` (a choice among the reference's kernels), and what the library
# returns FCC_ERR_UNSUPPORTED for (more than 4 body types or 4 body terms,
# nθ > 8000, more than 31 shape modes). `T = :mixed` / `:mixed_bf16` have no
# reference counterpart; the reference path then runs in Float32.
# `HIPEnsemble` runs independent replicas by one launch per kernel and step;
# `observer_start!` / `run_observed!` / `observations` accumulate long-run
# observables on the device.
module FourierContourCellsHIPExt

using FourierContourCells
using FourierContourCells: Bodies, PeriodicBox, BodyType, EnergyTerm, OnTypes, Tension,
    PerimeterElasticity, Overlap, LinearRepulsion, HarmonicBond, is_body_term, is_contact_term, is_bond_term,
    check_parents, nbodies, to_host
import FourierContourCells: step!, run!, compute_gradients!, update_neighbors!, shape_index,
    perimeters, areas, positions, unwrapped_positions, body_energy, contours
using FCCHIP: HIPBackend, libfcc_hip, Bending
import FCCHIP: HIPEnsemble, autotune!, observer_start!, observe!, run_observed!, observations,
    frame_request!, frame_read, checkpoint, restore!

const FCC_OK = Cint(0)
const FCC_ERR_ARGUMENT = Cint(-1)
const FCC_ERR_DIMENSION = Cint(-2)
const FCC_ERR_UNSUPPORTED = Cint(-3)
const FCC_FLAG_DEBUG_BUFFERS = Int32(1)
const PRECISION = Dict{Any, Int32}(Float64 => 0, Float32 => 1, :mixed => 2, :mixed_fp16 => 2, :mixed_bf16 => 3)

# ---- C structs (field for field as in include/fcc.h)

struct FccBodyType
    v0::Cdouble; Dr::Cdouble; Dt::Cdouble; mur::Cdouble; mutheta::Cdouble; muab::Cdouble
    parent_drag::Int32; reserved::Int32
end

struct FccTerm
    kind::Int32; n_on::Int32; on_types::NTuple{4, Int32}; p::NTuple{2, Cdouble}
    J::NTuple{16, Cdouble}; Jrel::NTuple{16, Cdouble}
end

struct FccConfig
    n_bodies::Int64; n_modes::Int32; n_theta::Int32; n_types::Int32; n_terms::Int32
    types::NTuple{4, FccBodyType}; terms::NTuple{8, FccTerm}
    dt::Cdouble; Lx::Cdouble; Ly::Cdouble; skin::Cdouble; seed::UInt64
    precision::Int32; flags::Int32
end

struct FccBodies
    x::Ptr{Cdouble}; y::Ptr{Cdouble}; theta::Ptr{Cdouble}; imgx::Ptr{Int32}; imgy::Ptr{Int32}
    coef::Ptr{Cdouble}; R0::Ptr{Cdouble}; type::Ptr{Int32}; parent::Ptr{Int32}
end

lasterror() = unsafe_string(ccall((:fcc_last_error, libfcc_hip), Cstring, ()))

# The library's error codes as the reference's exceptions (ArgumentError /
# DimensionMismatch for the host checks, src/simulation.jl:71-118).
function check(rc)
    rc == FCC_OK && return nothing
    msg = lasterror()
    rc == FCC_ERR_ARGUMENT && throw(ArgumentError(msg))
    rc == FCC_ERR_DIMENSION && throw(DimensionMismatch(msg))
    error("libfcc_hip: ", msg)
end

# ---- term and type conversion

const EMPTY_TERM = FccTerm(0, 0, (0, 0, 0, 0), (0.0, 0.0), ntuple(_ -> 0.0, 16), ntuple(_ -> 0.0, 16))

unwrap(t::OnTypes) = (t.term, Tuple(Int32.(t.types)))
unwrap(t) = (t, ())

table(J, nt) = ntuple(16) do q
    a, b = divrem(q - 1, 4)
    size(J, 1) == 1 ? J[1, 1] : (a < nt && b < nt ? J[a + 1, b + 1] : 0.0)
end

# The native form of a term, or `nothing` (no native kernel: reference path).
function fcc_term(t)
    base, on = unwrap(t)
    base isa OnTypes && return nothing  # nested OnTypes
    ont = ntuple(q -> q <= length(on) ? on[q] : Int32(0), 4)
    z = ntuple(_ -> 0.0, 16)
    if base isa Tension
        return FccTerm(1, length(on), ont, (Float64(base.η), 0.0), z, z)
    elseif base isa PerimeterElasticity
        return FccTerm(2, length(on), ont, (Float64(base.λ), Float64(base.p0)), z, z)
    elseif base isa Bending
        return FccTerm(3, length(on), ont, (Float64(base.κb), 0.0), z, z)
    elseif base isa Overlap
        nt = size(base.J, 1)
        return FccTerm(16, length(on), ont, (0.0, 0.0), table(Float64.(base.J), nt), table(Float64.(base.Jrelated), nt))
    elseif base isa LinearRepulsion
        return FccTerm(17, length(on), ont, (Float64(base.k), 0.0), z, z)
    elseif base isa HarmonicBond
        return FccTerm(32, length(on), ont, (Float64(base.k), Float64(base.r0)), z, z)
    end
    return nothing
end

fcc_type(b::BodyType) = FccBodyType(b.v0, b.Dr, b.Dt, b.μr, b.μθ, b.μab, b.parent_drag ? 1 : 0, 0)

host64(bodies) = (h = to_host(bodies);
    Bodies(map(f -> f isa AbstractArray{<:AbstractFloat} ? Float64.(f) : copy(f),
        (getfield(h, f) for f in fieldnames(Bodies)))...))

function fcc_config(host, box, types, terms, dt, nθ, skin, seed, T, flags)
    K = size(host.coef, 1) ÷ 2
    tt = ntuple(q -> q <= length(types) ? fcc_type(types[q]) : fcc_type(BodyType()), 4)
    te = ntuple(q -> q <= length(terms) ? terms[q] : EMPTY_TERM, 8)
    return FccConfig(nbodies(host), K, nθ, length(types), length(terms), tt, te, Float64(dt), Float64(box.Lx),
        Float64(box.Ly), Float64(skin), UInt64(seed), PRECISION[T], flags)
end

# ---- the reference path

# The reference's own Simulation (src/simulation.jl:66-105) on its CPU backend,
# with every argument the caller gave. Symbolic (mixed) precisions run in
# Float32 there.
function reference_simulation(why, bodies, box; T, kwargs...)
    Tref = T isa Type ? T : Float32
    @warn "FourierContourCells HIP backend: $why; using the reference path (T = $Tref)" maxlog = 1
    return FourierContourCells.Simulation(bodies, box; T = Tref, kwargs...)
end

# ---- simulation handle

"""
Host view of the gradients (`sim.grad`, src/gradients.jl:15-27): gx, gy, gθ,
Ebody [N]; glab, gcoef [ncoef × N]; with `debug = true` also P, GR, GD, GDD
[nθ × N] and Gall = [GR; GD; GDD] (otherwise `nothing`).
"""
struct HostGradients
    Gall::Union{Matrix{Float64}, Nothing}
    GR::Union{Matrix{Float64}, Nothing}
    GD::Union{Matrix{Float64}, Nothing}
    GDD::Union{Matrix{Float64}, Nothing}
    P::Union{Matrix{Float64}, Nothing}
    gx::Vector{Float64}
    gy::Vector{Float64}
    gθ::Vector{Float64}
    glab::Matrix{Float64}
    gcoef::Matrix{Float64}
    Ebody::Vector{Float64}
end

mutable struct HIPSimulation{T}
    handle::Ptr{Cvoid}
    host::Bodies               # host mirror of the body state (copied in/out)
    box::PeriodicBox{Float64}
    types::Tuple
    energies::Tuple
    nθ::Int
    debug::Bool
    dt::Float64
end

"""
    Simulation(bodies, box, ::HIPBackend; types, energies, bonds = (), dt = 0.005,
               nθ = 40, T = Float64, seed = 1, skin = 0.3, hook! = nothing,
               layout = nothing, debug = false)

`FourierContourCells.Simulation` (src/simulation.jl:66-105) on the MI355X.
`T` is Float64, Float32, `:mixed` (FP16 hi+lo matrix-core contractions) or
`:mixed_bf16`. `debug = true` keeps the contour gradients P, GR, GD, GDD for
`sim.grad`. Fields read like the reference's: `sim.bodies` (host copy of the
current state), `sim.t`, `sim.step`, `sim.dt` (settable), `sim.grad` (after
`compute_gradients!`), `sim.box`, `sim.types`.
"""
function FourierContourCells.Simulation(bodies::Bodies, box::PeriodicBox, ::HIPBackend; types,
        energies = (), bonds = (), dt = 0.005, nθ::Integer = 40, T = Float64, seed::Integer = 1,
        skin::Real = 0.3, hook! = nothing, layout = nothing, debug::Bool = false, kwargs...)
    energies = Tuple(energies)
    ref(why) = reference_simulation(why, bodies, box; T, types, energies, bonds, dt, nθ, seed, skin, hook!,
        (layout === nothing ? (;) : (; layout))..., kwargs...)
    isempty(kwargs) || return ref("keyword arguments $(Tuple(keys(kwargs))) are not native")
    hook! === nothing || return ref("hook! runs Julia code between the kernels of every step")
    layout === nothing || return ref("an explicit kernel layout selects the reference's kernels")
    haskey(PRECISION, T) || throw(ArgumentError("T must be Float64, Float32, :mixed or :mixed_bf16, got $T"))
    types64 = map(BodyType{Float64}, Tuple(types))
    all(t -> is_body_term(t) || is_contact_term(t) || is_bond_term(t), energies) ||
        throw(ArgumentError("energy terms must be BodyTerms, ContactTerms or BondTerms"))
    host = host64(bodies)
    maximum(host.type) <= length(types64) || throw(ArgumentError("body type without BodyType entry"))
    check_parents(host, types64)
    terms = map(fcc_term, energies)
    all(!isnothing, terms) || return ref("a term without a native kernel")
    (length(types64) <= 4 && length(terms) <= 8) || return ref("more than 4 body types or 8 terms")
    cfg = Ref(fcc_config(host, box, types64, terms, dt, nθ, skin, seed, T, debug ? FCC_FLAG_DEBUG_BUFFERS : Int32(0)))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = GC.@preserve host begin
        b = Ref(bodies_struct(host))
        ccall((:fcc_create, libfcc_hip), Cint, (Ref{FccConfig}, Ref{FccBodies}, Ref{Ptr{Cvoid}}), cfg, b, h)
    end
    rc == FCC_ERR_UNSUPPORTED && return ref(lasterror())
    check(rc)
    sim = HIPSimulation{T}(h[], host, PeriodicBox{Float64}(box.Lx, box.Ly), types64, energies, nθ, debug,
        Float64(dt))
    finalizer(s -> ccall((:fcc_destroy, libfcc_hip), Cint, (Ptr{Cvoid},), getfield(s, :handle)), sim)
    set_bonds!(sim, bonds)
    return sim
end

handle(sim::HIPSimulation) = getfield(sim, :handle)

bodies_struct(b::Bodies) = FccBodies(pointer(b.x), pointer(b.y), pointer(b.θ), pointer(b.imgx),
    pointer(b.imgy), pointer(b.coef), pointer(b.R0), pointer(b.type), pointer(b.parent))

function sync_host!(sim::HIPSimulation)
    host = getfield(sim, :host)
    GC.@preserve host begin
        b = Ref(bodies_struct(host))
        check(ccall((:fcc_get_bodies, libfcc_hip), Cint, (Ptr{Cvoid}, Ref{FccBodies}), handle(sim), b))
    end
    return host
end

function counters(sim::HIPSimulation)
    step, clock, builds, t = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0), Ref{Cdouble}(0.0)
    check(ccall((:fcc_get_counters, libfcc_hip), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Cdouble}, Ref{Int64},
        Ref{Int64}), handle(sim), step, t, clock, builds))
    return (; step = Int(step[]), t = t[], clock = clock[], builds = Int(builds[]))
end

function read_gradients(sim::HIPSimulation)
    N, nc, nθ = nbodies(sim), size(getfield(sim, :host).coef, 1), getfield(sim, :nθ)
    gx, gy, gθ, E = zeros(N), zeros(N), zeros(N), zeros(N)
    glab, gcoef = zeros(nc, N), zeros(nc, N)
    check(ccall((:fcc_get_gradients, libfcc_hip), Cint,
        (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
        handle(sim), gx, gy, gθ, glab, gcoef, E))
    P = GR = GD = GDD = Gall = nothing
    if getfield(sim, :debug)
        P, GR, GD, GDD = zeros(nθ, N), zeros(nθ, N), zeros(nθ, N), zeros(nθ, N)
        check(ccall((:fcc_get_contour_gradients, libfcc_hip), Cint,
            (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}), handle(sim), P, GR, GD, GDD))
        Gall = vcat(GR, GD, GDD)
    end
    return HostGradients(Gall, GR, GD, GDD, P, gx, gy, gθ, glab, gcoef, E)
end

function read_geometry(sim::HIPSimulation)
    N, nc, nθ = nbodies(sim), size(getfield(sim, :host).coef, 1), getfield(sim, :nθ)
    lab, rmax, R, dR, ddR = zeros(nc, N), zeros(N), zeros(nθ, N), zeros(nθ, N), zeros(nθ, N)
    check(ccall((:fcc_get_geometry, libfcc_hip), Cint,
        (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
        handle(sim), lab, rmax, R, dR, ddR))
    return (; lab, rmax, R, dR, ddR)
end

# The reference's fields (src/simulation.jl:38-64) as properties.
function Base.getproperty(sim::HIPSimulation, s::Symbol)
    s === :bodies && return sync_host!(sim)
    s === :t && return counters(sim).t
    s === :step && return counters(sim).step
    s === :grad && return read_gradients(sim)
    s === :geom && return read_geometry(sim)
    s === :clock && return [counters(sim).clock]
    return getfield(sim, s)
end
Base.propertynames(::HIPSimulation) = (:bodies, :box, :types, :energies, :nθ, :t, :step, :dt, :grad, :geom, :clock)

# sim.dt = v, as the reference's mutable field (src/simulation.jl:61)
function Base.setproperty!(sim::HIPSimulation, s::Symbol, v)
    s === :dt || error("HIPSimulation: only dt can be set")
    check(ccall((:fcc_set_dt, libfcc_hip), Cint, (Ptr{Cvoid}, Cdouble), handle(sim), Float64(v)))
    setfield!(sim, :dt, Float64(v))
    return v
end

FourierContourCells.nbodies(sim::HIPSimulation) = nbodies(getfield(sim, :host))

function Base.show(io::IO, sim::HIPSimulation{T}) where {T}
    c = counters(sim)
    print(io, "Simulation{$T}(N=$(nbodies(sim)), nθ=$(getfield(sim, :nθ)), box=$(sim.box.Lx)×$(sim.box.Ly), ",
        "t=$(c.t), backend=HIPBackend())")
end

# step!(sim, n) — src/simulation.jl:165-173
step!(sim::HIPSimulation, n::Integer = 1) =
    (check(ccall((:fcc_step, libfcc_hip), Cint, (Ptr{Cvoid}, Int64), handle(sim), n)); sim)

# run!(sim, n; every, callback, graph) — src/simulation.jl:220-235 with the
# graph replay of ext/FourierContourCellsCUDAExt.jl:144-180
function run!(sim::HIPSimulation, nsteps::Integer; every::Integer = 0, callback = nothing, graph::Bool = true)
    adv(n) = graph ? check(ccall((:fcc_run, libfcc_hip), Cint, (Ptr{Cvoid}, Int64), handle(sim), n)) :
        step!(sim, n)
    if callback === nothing || every <= 0
        adv(nsteps)
    else
        done = 0
        while done < nsteps
            n = min(every, nsteps - done)
            adv(n)
            done += n
            callback(sim)
        end
    end
    return sim
end

# update_neighbors!(sim; force) — src/simulation.jl:132-142
update_neighbors!(sim::HIPSimulation; force::Bool = false) =
    (check(ccall((:fcc_update_neighbors, libfcc_hip), Cint, (Ptr{Cvoid}, Int32), handle(sim), force)); sim)

# compute_gradients!(sim; frame) — src/simulation.jl:150-158; returns the
# gradients like the reference (which returns sim.grad)
function compute_gradients!(sim::HIPSimulation; frame::Bool = true)
    check(ccall((:fcc_compute_gradients, libfcc_hip), Cint, (Ptr{Cvoid}, Int32), handle(sim), frame))
    return read_gradients(sim)
end

function observables(sim::HIPSimulation)
    N = nbodies(sim)
    l, a, p = zeros(N), zeros(N), zeros(N)
    check(ccall((:fcc_shape_index, libfcc_hip), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
        handle(sim), l, a, p))
    return l, a, p
end

select(v, sim, ::Nothing) = v
select(v, sim, type::Integer) = v[getfield(sim, :host).type .== type]

# src/observables.jl:6-18
perimeters(sim::HIPSimulation; type = nothing) = select(observables(sim)[1], sim, type)
areas(sim::HIPSimulation; type = nothing) = select(observables(sim)[2], sim, type)
shape_index(sim::HIPSimulation; type = nothing) = select(observables(sim)[3], sim, type)

# src/observables.jl:21-29
positions(sim::HIPSimulation) = (b = sync_host!(sim); permutedims(hcat(b.x, b.y)))
function unwrapped_positions(sim::HIPSimulation)
    b, box = sync_host!(sim), getfield(sim, :box)
    return permutedims(hcat(b.x .+ b.imgx .* box.Lx, b.y .+ b.imgy .* box.Ly))
end
body_energy(sim::HIPSimulation) = sum(compute_gradients!(sim).Ebody)

# src/observables.jl:36-44
function contours(sim::HIPSimulation)
    X, Y = zeros(getfield(sim, :nθ), nbodies(sim)), zeros(getfield(sim, :nθ), nbodies(sim))
    check(ccall((:fcc_contours, libfcc_hip), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}), handle(sim), X, Y))
    return X, Y
end

# frames for live views: the copy of frame k overlaps the steps issued after it
frame_request!(sim::HIPSimulation) =
    (check(ccall((:fcc_frame_request, libfcc_hip), Cint, (Ptr{Cvoid},), handle(sim))); sim)
function frame_read(sim::HIPSimulation)
    X, Y = zeros(getfield(sim, :nθ), nbodies(sim)), zeros(getfield(sim, :nθ), nbodies(sim))
    clock = Ref{Int64}(0)
    check(ccall((:fcc_frame_read, libfcc_hip), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ref{Int64}),
        handle(sim), X, Y, clock))
    return X, Y, clock[]
end

# checkpoint / resume (bitwise continuation, with the device clock)
function checkpoint(sim::HIPSimulation)
    n = Ref{Int64}(0)
    check(ccall((:fcc_checkpoint_save, libfcc_hip), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Int64, Ref{Int64}),
        handle(sim), C_NULL, 0, n))
    blob = Vector{UInt8}(undef, n[])
    check(ccall((:fcc_checkpoint_save, libfcc_hip), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Int64, Ref{Int64}),
        handle(sim), blob, length(blob), n))
    return blob
end
restore!(sim::HIPSimulation, blob::Vector{UInt8}) =
    (check(ccall((:fcc_checkpoint_load, libfcc_hip), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Int64), handle(sim), blob,
        length(blob))); sim)

# `bonds` of Simulation (src/simulation.jl:14-15): pairs (i, j), 1-based,
# checked and built into the bond list like BondList (src/energies/bonds.jl:54-72)
function set_bonds!(sim::HIPSimulation, bonds)
    pairs = Int32[v for b in bonds for v in (b[1], b[2])]
    check(ccall((:fcc_set_bonds, libfcc_hip), Cint, (Ptr{Cvoid}, Int64, Ptr{Int32}), handle(sim),
        length(pairs) ÷ 2, pairs))
    return sim
end

# ---- long-run observables on the device: run!(sim, n; every, callback) with a
# callback evaluating shape_index / unwrapped_positions (src/observables.jl:18-29)

struct FccObserverConfig
    nbins::Int32; max_samples::Int32; pmin::Cdouble; pmax::Cdouble
end

const OBSERVERS = IdDict{Any, FccObserverConfig}()

function observer_start!(sim::HIPSimulation; nbins::Integer = 256, prange = (3.5, 5.5), max_samples::Integer = 1024)
    c = FccObserverConfig(nbins, max_samples, prange[1], prange[2])
    check(ccall((:fcc_observer_start, libfcc_hip), Cint, (Ptr{Cvoid}, Ref{FccObserverConfig}), handle(sim), c))
    OBSERVERS[sim] = c
    return sim
end

observe!(sim::HIPSimulation) = (check(ccall((:fcc_observe, libfcc_hip), Cint, (Ptr{Cvoid},), handle(sim))); sim)

run_observed!(sim::HIPSimulation, nsteps::Integer; every::Integer) =
    (check(ccall((:fcc_run_observed, libfcc_hip), Cint, (Ptr{Cvoid}, Int64, Int64), handle(sim), nsteps, every)); sim)

"Samples so far: clock, msd / p_mean / p2_mean (types x samples), hist ((nbins + 2) x types), edges."
function observations(sim::HIPSimulation)
    c = OBSERVERS[sim]
    nt, ms = length(getfield(sim, :types)), Int(c.max_samples)
    ns = Ref{Int32}(0)
    clock = zeros(Int64, ms)
    msd, pm, p2 = zeros(nt, ms), zeros(nt, ms), zeros(nt, ms)
    hist = zeros(UInt64, c.nbins + 2, nt)
    check(ccall((:fcc_observer_read, libfcc_hip), Cint,
        (Ptr{Cvoid}, Ref{Int32}, Ptr{Int64}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{UInt64}),
        handle(sim), ns, clock, msd, pm, p2, hist))
    n = Int(ns[])
    edges = c.pmin .+ (c.pmax - c.pmin) .* (0:c.nbins) ./ c.nbins
    return (; clock = clock[1:n], msd = msd[:, 1:n], p_mean = pm[:, 1:n], p2_mean = p2[:, 1:n], hist, edges)
end

# ---- ensembles: independent simulations advanced by one launch per kernel
# and step (README.md:249-261 runs them as concurrent tasks instead)

struct FccLaunch
    update_block::Int32; xcd_order::Int32; chunk::Int32; lanes::Int32; kernels::Int32
end

mutable struct HIPEnsembleHandle
    handle::Ptr{Cvoid}
    members::Vector{HIPSimulation}
end

"""
    HIPEnsemble(specs; T = Float32, nθ = 40)

`specs` is a vector of named tuples `(bodies, box, types, energies, dt = 0.005,
seed = 1, skin = 0.3, bonds = ())`, one per replica. Replicas share `T`, `nθ` and the mode
count. `members` are `HIPSimulation`s (borrowed handles).
"""
function HIPEnsemble(specs::AbstractVector; T = Float32, nθ::Integer = 40, debug::Bool = false)
    haskey(PRECISION, T) || throw(ArgumentError("T must be Float64, Float32, :mixed or :mixed_bf16, got $T"))
    hosts, cfgs, typesets, termsets = Bodies[], FccConfig[], Tuple[], Tuple[]
    flags = debug ? FCC_FLAG_DEBUG_BUFFERS : Int32(0)
    for sp in specs
        types = map(BodyType{Float64}, Tuple(sp.types))
        energies = Tuple(sp.energies)
        terms = map(fcc_term, energies)
        all(!isnothing, terms) || throw(ArgumentError("ensembles need energy terms with native kernels"))
        host = host64(sp.bodies)
        maximum(host.type) <= length(types) || throw(ArgumentError("body type without BodyType entry"))
        check_parents(host, types)
        push!(cfgs, fcc_config(host, sp.box, types, terms, get(sp, :dt, 0.005), nθ, get(sp, :skin, 0.3),
            get(sp, :seed, 1), T, flags))
        push!(hosts, host)
        push!(typesets, types)
        push!(termsets, energies)
    end
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve hosts begin
        bs = [bodies_struct(b) for b in hosts]
        check(ccall((:fcc_ensemble_create, libfcc_hip), Cint,
            (Int32, Ptr{FccConfig}, Ptr{FccBodies}, Ref{Ptr{Cvoid}}), length(cfgs), cfgs, bs, h))
    end
    members = HIPSimulation[]
    for (r, sp) in enumerate(specs)
        m = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:fcc_ensemble_member, libfcc_hip), Cint, (Ptr{Cvoid}, Int32, Ref{Ptr{Cvoid}}),
            h[], r - 1, m))
        push!(members, HIPSimulation{T}(m[], hosts[r], PeriodicBox{Float64}(sp.box.Lx, sp.box.Ly), typesets[r],
            termsets[r], nθ, debug, Float64(get(sp, :dt, 0.005))))
        set_bonds!(members[end], get(sp, :bonds, ()))
    end
    ens = HIPEnsembleHandle(h[], members)
    finalizer(e -> ccall((:fcc_ensemble_destroy, libfcc_hip), Cint, (Ptr{Cvoid},), e.handle), ens)
    return ens
end

# run!(sim, n) of every replica
run!(e::HIPEnsembleHandle, nsteps::Integer) =
    (check(ccall((:fcc_ensemble_run, libfcc_hip), Cint, (Ptr{Cvoid}, Int64), e.handle, nsteps)); e)

function observer_start!(e::HIPEnsembleHandle; nbins::Integer = 256, prange = (3.5, 5.5),
        max_samples::Integer = 1024)
    c = FccObserverConfig(nbins, max_samples, prange[1], prange[2])
    check(ccall((:fcc_ensemble_observer_start, libfcc_hip), Cint, (Ptr{Cvoid}, Ref{FccObserverConfig}),
        e.handle, c))
    foreach(m -> OBSERVERS[m] = c, e.members)
    return e
end

observe!(e::HIPEnsembleHandle) = (check(ccall((:fcc_ensemble_observe, libfcc_hip), Cint, (Ptr{Cvoid},), e.handle)); e)

run_observed!(e::HIPEnsembleHandle, nsteps::Integer; every::Integer) =
    (check(ccall((:fcc_ensemble_run_observed, libfcc_hip), Cint, (Ptr{Cvoid}, Int64, Int64), e.handle, nsteps,
        every)); e)

function autotune!(e::HIPEnsembleHandle; steps::Integer = 20)
    best = Ref(FccLaunch(0, 0, 0, 0, 0))
    ms = Ref{Cdouble}(0.0)
    check(ccall((:fcc_ensemble_autotune, libfcc_hip), Cint, (Ptr{Cvoid}, Int64, Ref{FccLaunch}, Ref{Cdouble}),
        e.handle, steps, best, ms))
    return best[], ms[]
end

end
