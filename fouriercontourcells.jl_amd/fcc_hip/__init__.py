"""fcc_hip — host mirror of the FourierContourCells.jl time-step API over
libfcc_hip.so (MI355X / gfx950)."""
from .api import (ArgumentError, Bending, BodyType, Bodies, DeviceError, DimensionMismatch, Ensemble,
                  HarmonicBond, LinearRepulsion, OnTypes, Overlap, PerimeterElasticity, PeriodicBox,
                  Replica, Simulation, Tension, UnsupportedError, areas, body_energy, compute_gradients,
                  contours, frame_read, frame_request, hexagonal_monolayer, observations, observe,
                  observer_start, parent_bonds, perimeters, philox_noise, philox_uniforms, positions,
                  refresh, run, run_observed, shape_index, step, unwrapped_positions, update_neighbors,
                  vcat, with_nuclei)
from ._abi import hip_engine, library_path
