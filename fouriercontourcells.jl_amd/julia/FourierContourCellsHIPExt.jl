# FourierContourCellsHIPExt — package extension of FourierContourCells.jl that
# runs the time step on an AMD MI355X through libfcc_hip.so (C ABI of
# include/fcc.h). Loaded through the reference's weak-dependency mechanism
# (Project.toml [weakdeps]/[extensions], like FourierContourCellsCUDAExt): the
# weak dependency `FCCHIP` is a tiny package that provides `HIPBackend` and the
# path of the built library.
#
#   using FourierContourCells, FCCHIP
#   sim = Simulation(bodies, box, HIPBackend(); types, energies, dt, T = Float32)
#   run!(sim, 10_000)
#   p = shape_index(sim)
#
# Terms or sizes without a native kernel (user BodyTerms and BondTerms,
# nθ > 64, more than 4 body types) make `fcc_create` return
# FCC_ERR_UNSUPPORTED; the extension then builds the reference Simulation on
# the CPU backend instead (a user-defined term runs on the reference path).
# `HIPEnsemble` runs independent replicas by one launch per kernel and step;
# `observer_start!` / `run_observed!` / `observations` accumulate long-run
# observables on the device.
module FourierContourCellsHIPExt

using FourierContourCells
using FourierContourCells: Bodies, PeriodicBox, BodyType, EnergyTerm, OnTypes, Tension,
    PerimeterElasticity, Overlap, LinearRepulsion, HarmonicBond, is_body_term, is_contact_term, is_bond_term, check_parents,
    nbodies, to_host
import FourierContourCells: step!, run!, compute_gradients!, update_neighbors!, shape_index,
    perimeters, areas, positions, unwrapped_positions, body_energy, contours
using FCCHIP: HIPBackend, libfcc_hip, Bending
import FCCHIP: HIPEnsemble, autotune!, observer_start!, observe!, run_observed!, observations,
    frame_request!, frame_read, checkpoint, restore!

const FCC_OK = Cint(0)
const FCC_ERR_UNSUPPORTED = Cint(-3)
const PRECISION = Dict(Float64 => Int32(0), Float32 => Int32(1), :mixed => Int32(2), :mixed_bf16 => Int32(3))

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
check(rc) = rc == FCC_OK || error("libfcc_hip: ", lasterror())

# ---- term and type conversion

const EMPTY_TERM = FccTerm(0, 0, (0, 0, 0, 0), (0.0, 0.0), ntuple(_ -> 0.0, 16), ntuple(_ -> 0.0, 16))

unwrap(t::OnTypes) = (t.term, Tuple(Int32.(t.types)))
unwrap(t) = (t, ())

table(J, nt) = ntuple(16) do q
    a, b = divrem(q - 1, 4)
    size(J, 1) == 1 ? J[1, 1] : (a < nt && b < nt ? J[a + 1, b + 1] : 0.0)
end

function fcc_term(t)
    base, on = unwrap(t)
    ont = ntuple(q -> q <= length(on) ? on[q] : Int32(0), 4)
    if base isa Tension
        return FccTerm(1, length(on), ont, (base.η, 0.0), ntuple(_ -> 0.0, 16), ntuple(_ -> 0.0, 16))
    elseif base isa PerimeterElasticity
        return FccTerm(2, length(on), ont, (base.λ, base.p0), ntuple(_ -> 0.0, 16), ntuple(_ -> 0.0, 16))
    elseif base isa Bending
        return FccTerm(3, length(on), ont, (base.κb, 0.0), ntuple(_ -> 0.0, 16), ntuple(_ -> 0.0, 16))
    elseif base isa Overlap
        nt = size(base.J, 1)
        return FccTerm(16, length(on), ont, (0.0, 0.0), table(Float64.(base.J), nt), table(Float64.(base.Jrelated), nt))
    elseif base isa LinearRepulsion
        return FccTerm(17, length(on), ont, (Float64(base.k), 0.0), ntuple(_ -> 0.0, 16), ntuple(_ -> 0.0, 16))
    elseif base isa HarmonicBond
        return FccTerm(32, length(on), ont, (Float64(base.k), Float64(base.r0)), ntuple(_ -> 0.0, 16),
            ntuple(_ -> 0.0, 16))
    end
    return nothing  # no native kernel
end

fcc_type(b::BodyType) = FccBodyType(b.v0, b.Dr, b.Dt, b.μr, b.μθ, b.μab, b.parent_drag ? 1 : 0, 0)

# ---- simulation handle

mutable struct HIPSimulation{T}
    handle::Ptr{Cvoid}
    host::Bodies               # host mirror of the body state (copied in/out)
    box::PeriodicBox{Float64}
    nθ::Int
    dt::Float64
    ntypes::Int
end

"""
    Simulation(bodies, box, ::HIPBackend; types, energies, dt = 0.005, nθ = 40,
               T = Float64, seed = 1, skin = 0.3)

`FourierContourCells.Simulation` on the MI355X. `T` is Float64, Float32,
`:mixed` (FP16 hi+lo matrix-core contractions) or `:mixed_bf16`.
"""
function FourierContourCells.Simulation(bodies::Bodies, box::PeriodicBox, ::HIPBackend; types,
        energies = (), bonds = (), dt = 0.005, nθ::Integer = 40, T = Float64, seed::Integer = 1,
        skin::Real = 0.3, kwargs...)
    types = map(BodyType{Float64}, Tuple(types))
    energies = Tuple(energies)
    all(t -> is_body_term(t) || is_contact_term(t) || is_bond_term(t), energies) ||
        throw(ArgumentError("energy terms must be BodyTerms, ContactTerms or BondTerms"))
    host = to_host(bodies)
    maximum(host.type) <= length(types) || throw(ArgumentError("body type without BodyType entry"))
    host = Bodies(map(f -> f isa AbstractArray{<:AbstractFloat} ? Float64.(f) : f,
        (getfield(host, f) for f in fieldnames(Bodies)))...)
    check_parents(host, types)
    terms = map(fcc_term, energies)
    native = all(!isnothing, terms) && length(types) <= 4 && length(terms) <= 8
    if !native
        @warn "no native kernel for this configuration: using the reference path" maxlog = 1
        return FourierContourCells.Simulation(bodies, box; types, energies, bonds, dt, nθ,
            T = T isa Type ? T : Float32, seed, skin, kwargs...)
    end
    K = size(host.coef, 1) ÷ 2
    tt = ntuple(q -> q <= length(types) ? fcc_type(types[q]) : fcc_type(BodyType()), 4)
    te = ntuple(q -> q <= length(terms) ? terms[q] : EMPTY_TERM, 8)
    cfg = Ref(FccConfig(nbodies(host), K, nθ, length(types), length(terms), tt, te, dt, box.Lx, box.Ly,
        skin, UInt64(seed), PRECISION[T], 0))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve host begin
        b = Ref(bodies_struct(host))
        rc = ccall((:fcc_create, libfcc_hip), Cint, (Ref{FccConfig}, Ref{FccBodies}, Ref{Ptr{Cvoid}}),
            cfg, b, h)
    end
    if rc == FCC_ERR_UNSUPPORTED
        @warn "libfcc_hip: $(lasterror()); using the reference path" maxlog = 1
        return FourierContourCells.Simulation(bodies, box; types, energies, bonds, dt, nθ, seed, skin)
    end
    check(rc)
    sim = HIPSimulation{T}(h[], host, PeriodicBox{Float64}(box.Lx, box.Ly), nθ, dt, length(types))
    finalizer(s -> ccall((:fcc_destroy, libfcc_hip), Cint, (Ptr{Cvoid},), s.handle), sim)
    set_bonds!(sim, bonds)
    return sim
end

bodies_struct(b::Bodies) = FccBodies(pointer(b.x), pointer(b.y), pointer(b.θ), pointer(b.imgx),
    pointer(b.imgy), pointer(b.coef), pointer(b.R0), pointer(b.type), pointer(b.parent))

function sync_host!(sim::HIPSimulation)
    GC.@preserve sim begin
        b = Ref(bodies_struct(sim.host))
        check(ccall((:fcc_get_bodies, libfcc_hip), Cint, (Ptr{Cvoid}, Ref{FccBodies}), sim.handle, b))
    end
    return sim.host
end

FourierContourCells.nbodies(sim::HIPSimulation) = nbodies(sim.host)

# step!(sim, n) — src/simulation.jl:165-173
step!(sim::HIPSimulation, n::Integer = 1) =
    (check(ccall((:fcc_step, libfcc_hip), Cint, (Ptr{Cvoid}, Int64), sim.handle, n)); sim)

# run!(sim, n; every, callback, graph) — src/simulation.jl:220-235 with the
# graph replay of ext/FourierContourCellsCUDAExt.jl:144-180
function run!(sim::HIPSimulation, nsteps::Integer; every::Integer = 0, callback = nothing, graph::Bool = true)
    adv(n) = graph ? check(ccall((:fcc_run, libfcc_hip), Cint, (Ptr{Cvoid}, Int64), sim.handle, n)) : step!(sim, n)
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

update_neighbors!(sim::HIPSimulation; force::Bool = false) =
    (check(ccall((:fcc_update_neighbors, libfcc_hip), Cint, (Ptr{Cvoid}, Int32), sim.handle, force)); sim)

function compute_gradients!(sim::HIPSimulation; frame::Bool = true)
    check(ccall((:fcc_compute_gradients, libfcc_hip), Cint, (Ptr{Cvoid}, Int32), sim.handle, frame))
    N, nc = nbodies(sim), size(sim.host.coef, 1)
    gx, gy, gθ, E = zeros(N), zeros(N), zeros(N), zeros(N)
    glab, gcoef = zeros(nc, N), zeros(nc, N)
    check(ccall((:fcc_get_gradients, libfcc_hip), Cint,
        (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
        sim.handle, gx, gy, gθ, glab, gcoef, E))
    return (; gx, gy, gθ, glab, gcoef, Ebody = E)
end

function observables(sim::HIPSimulation)
    N = nbodies(sim)
    l, a, p = zeros(N), zeros(N), zeros(N)
    check(ccall((:fcc_shape_index, libfcc_hip), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
        sim.handle, l, a, p))
    return l, a, p
end

select(v, sim, ::Nothing) = v
select(v, sim, type::Integer) = v[sim.host.type .== type]

# src/observables.jl:6-18
perimeters(sim::HIPSimulation; type = nothing) = select(observables(sim)[1], sim, type)
areas(sim::HIPSimulation; type = nothing) = select(observables(sim)[2], sim, type)
shape_index(sim::HIPSimulation; type = nothing) = select(observables(sim)[3], sim, type)

positions(sim::HIPSimulation) = (b = sync_host!(sim); permutedims(hcat(b.x, b.y)))
function unwrapped_positions(sim::HIPSimulation)
    b = sync_host!(sim)
    return permutedims(hcat(b.x .+ b.imgx .* sim.box.Lx, b.y .+ b.imgy .* sim.box.Ly))
end
body_energy(sim::HIPSimulation) = sum(compute_gradients!(sim).Ebody)

# src/observables.jl:36-44
function contours(sim::HIPSimulation)
    X, Y = zeros(sim.nθ, nbodies(sim)), zeros(sim.nθ, nbodies(sim))
    check(ccall((:fcc_contours, libfcc_hip), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}), sim.handle, X, Y))
    return X, Y
end

# frames for live views: the copy of frame k overlaps the steps issued after it
frame_request!(sim::HIPSimulation) =
    (check(ccall((:fcc_frame_request, libfcc_hip), Cint, (Ptr{Cvoid},), sim.handle)); sim)
function frame_read(sim::HIPSimulation)
    X, Y = zeros(sim.nθ, nbodies(sim)), zeros(sim.nθ, nbodies(sim))
    clock = Ref{Int64}(0)
    check(ccall((:fcc_frame_read, libfcc_hip), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ref{Int64}),
        sim.handle, X, Y, clock))
    return X, Y, clock[]
end

# checkpoint / resume (bitwise continuation, with the device clock)
function checkpoint(sim::HIPSimulation)
    n = Ref{Int64}(0)
    check(ccall((:fcc_checkpoint_save, libfcc_hip), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Int64, Ref{Int64}),
        sim.handle, C_NULL, 0, n))
    blob = Vector{UInt8}(undef, n[])
    check(ccall((:fcc_checkpoint_save, libfcc_hip), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Int64, Ref{Int64}),
        sim.handle, blob, length(blob), n))
    return blob
end
restore!(sim::HIPSimulation, blob::Vector{UInt8}) =
    (check(ccall((:fcc_checkpoint_load, libfcc_hip), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Int64), sim.handle, blob,
        length(blob))); sim)

# `bonds` of Simulation (src/simulation.jl:14-15): pairs (i, j), 1-based,
# checked and built into the bond list like BondList (src/energies/bonds.jl:54-72)
function set_bonds!(sim::HIPSimulation, bonds)
    pairs = Int32[v for b in bonds for v in (b[1], b[2])]
    check(ccall((:fcc_set_bonds, libfcc_hip), Cint, (Ptr{Cvoid}, Int64, Ptr{Int32}), sim.handle,
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
    check(ccall((:fcc_observer_start, libfcc_hip), Cint, (Ptr{Cvoid}, Ref{FccObserverConfig}), sim.handle, c))
    OBSERVERS[sim] = c
    return sim
end

observe!(sim::HIPSimulation) = (check(ccall((:fcc_observe, libfcc_hip), Cint, (Ptr{Cvoid},), sim.handle)); sim)

run_observed!(sim::HIPSimulation, nsteps::Integer; every::Integer) =
    (check(ccall((:fcc_run_observed, libfcc_hip), Cint, (Ptr{Cvoid}, Int64, Int64), sim.handle, nsteps, every)); sim)

"Samples so far: clock, msd / p_mean / p2_mean (types x samples), hist ((nbins + 2) x types), edges."
function observations(sim::HIPSimulation)
    c = OBSERVERS[sim]
    nt, ms = sim.ntypes, Int(c.max_samples)
    ns = Ref{Int32}(0)
    clock = zeros(Int64, ms)
    msd, pm, p2 = zeros(nt, ms), zeros(nt, ms), zeros(nt, ms)
    hist = zeros(UInt64, c.nbins + 2, nt)
    check(ccall((:fcc_observer_read, libfcc_hip), Cint,
        (Ptr{Cvoid}, Ref{Int32}, Ptr{Int64}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{UInt64}),
        sim.handle, ns, clock, msd, pm, p2, hist))
    n = Int(ns[])
    edges = c.pmin .+ (c.pmax - c.pmin) .* (0:c.nbins) ./ c.nbins
    return (; clock = clock[1:n], msd = msd[:, 1:n], p_mean = pm[:, 1:n], p2_mean = p2[:, 1:n], hist, edges)
end

# ---- ensembles: independent simulations advanced by one launch per kernel
# and step (README.md:249-261 runs them as concurrent tasks instead)

struct FccLaunch
    update_block::Int32; xcd_order::Int32; chunk::Int32; reserved::Int32
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
function HIPEnsemble(specs::AbstractVector; T = Float32, nθ::Integer = 40)
    hosts = Bodies[]
    cfgs = FccConfig[]
    for sp in specs
        types = map(BodyType{Float64}, Tuple(sp.types))
        terms = map(fcc_term, Tuple(sp.energies))
        all(!isnothing, terms) || throw(ArgumentError("ensembles need native energy terms"))
        host = to_host(sp.bodies)
        host = Bodies(map(f -> f isa AbstractArray{<:AbstractFloat} ? Float64.(f) : f,
            (getfield(host, f) for f in fieldnames(Bodies)))...)
        check_parents(host, types)
        K = size(host.coef, 1) ÷ 2
        tt = ntuple(q -> q <= length(types) ? fcc_type(types[q]) : fcc_type(BodyType()), 4)
        te = ntuple(q -> q <= length(terms) ? terms[q] : EMPTY_TERM, 8)
        push!(cfgs, FccConfig(nbodies(host), K, nθ, length(types), length(terms), tt, te,
            get(sp, :dt, 0.005), sp.box.Lx, sp.box.Ly, get(sp, :skin, 0.3), UInt64(get(sp, :seed, 1)),
            PRECISION[T], 0))
        push!(hosts, host)
    end
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve hosts begin
        bs = [bodies_struct(b) for b in hosts]
        check(ccall((:fcc_ensemble_create, libfcc_hip), Cint,
            (Int32, Ptr{FccConfig}, Ptr{FccBodies}, Ref{Ptr{Cvoid}}), length(cfgs), cfgs, bs, h))
    end
    members = HIPSimulation[]
    for (r, (sp, host)) in enumerate(zip(specs, hosts))
        m = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:fcc_ensemble_member, libfcc_hip), Cint, (Ptr{Cvoid}, Int32, Ref{Ptr{Cvoid}}),
            h[], r - 1, m))
        push!(members, HIPSimulation{T}(m[], host, PeriodicBox{Float64}(sp.box.Lx, sp.box.Ly), nθ,
            get(sp, :dt, 0.005), length(sp.types)))
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
    best = Ref(FccLaunch(0, 0, 0, 0))
    ms = Ref{Cdouble}(0.0)
    check(ccall((:fcc_ensemble_autotune, libfcc_hip), Cint, (Ptr{Cvoid}, Int64, Ref{FccLaunch}, Ref{Cdouble}),
        e.handle, steps, best, ms))
    return best[], ms[]
end

end
