"""
    FCCHIP

The weak dependency that switches on FourierContourCells.jl's MI355X backend.
`using FourierContourCells, FCCHIP` loads the package extension
`FourierContourCellsHIPExt`, which runs the time step through the C ABI of
`libfcc_hip.so` (include/fcc.h):

    sim = Simulation(bodies, box, HIPBackend(); types, energies, dt, T = :mixed)
    run!(sim, 10_000)
    shape_index(sim)

This package holds only what the extension needs to exist before it loads: the
backend selector `HIPBackend`, the built-in `Bending` term (a user term in
the reference, test/test_energies.jl:6-10, native here), the library path and
the generic functions of the HIP-only features, whose methods the extension
defines.
"""
module FCCHIP

using Libdl: Libdl
using FourierContourCells: FourierContourCells, BodyTerm
import FourierContourCells: contour_integrands, energy

export HIPBackend, Bending, HIPEnsemble, autotune!, observer_start!, observe!, run_observed!,
    observations, frame_request!, frame_read, checkpoint, restore!

"Selects the MI355X time step: `Simulation(bodies, box, HIPBackend(); ...)`."
struct HIPBackend end

"""
    Bending(κb)

Bending energy κb ∮ (R² + 2R'² − R R'')² / (R² + R'²)^{5/2} dφ, the term of
test/test_energies.jl:6-10 (examples/custom_energy.jl:42-49). The extension
evaluates it natively; on the reference path it is an ordinary user
`BodyTerm` (gradients by automatic differentiation).
"""
struct Bending{T} <: BodyTerm
    κb::T
end

function contour_integrands(::Bending, R, dR, ddR)
    s = sqrt(R^2 + dR^2)
    return (R^2 + 2dR^2 - R * ddR)^2 / (s^4 * s)
end
energy(t::Bending, Q, b) = t.κb * Q[1]

# The library built from fouriercontourcells.jl_amd/csrc (make), or the file
# named by ENV["FCC_HIP_LIBRARY"] when the package is precompiled.
const libfcc_hip = get(ENV, "FCC_HIP_LIBRARY",
    normpath(joinpath(@__DIR__, "..", "..", "..", "csrc", "libfcc_hip.so")))

# The C ABI version this package and the extension are written against
# (FCC_ABI_VERSION of include/fcc.h).
const ABI_VERSION = 4

function __init__()
    isfile(libfcc_hip) || @warn "libfcc_hip.so not found at $libfcc_hip: build it with make in csrc/"
    isfile(libfcc_hip) || return
    Libdl.dlopen(libfcc_hip)
    v = ccall((:fcc_abi_version, libfcc_hip), Cint, ())
    v == ABI_VERSION || error("libfcc_hip.so has C ABI version $v, FCCHIP expects $ABI_VERSION")
    return
end

# HIP-only features; the methods are in the extension.
"`HIPEnsemble(specs; T, nθ)`: independent replicas advanced by one launch per kernel and step."
function HIPEnsemble end
"`autotune!(ens; steps)`: the fastest launch shape of an ensemble step (state unchanged)."
function autotune! end
"`observer_start!(sim; nbins, prange, max_samples)`: start device-side long-run observables."
function observer_start! end
"`observe!(sim)`: one device sample of the current state."
function observe! end
"`run_observed!(sim, n; every)`: `run!(sim, n; every, callback)` with a device sample as the callback."
function run_observed! end
"`observations(sim)`: the device samples so far."
function observations end
"`frame_request!(sim)`: enqueue the contours of the current state for an asynchronous copy."
function frame_request! end
"`frame_read(sim)`: wait for the requested frame and return (X, Y, clock)."
function frame_read end
"`checkpoint(sim)`: the run state (with the device clock) as bytes."
function checkpoint end
"`restore!(sim, blob)`: continue bitwise from `checkpoint` bytes."
function restore! end

end
