/*
 * fcc_bench.h — measurement entry point of libfcc_hip.so.
 *
 * No counterpart in the reference (whose timings are wall-clock around run!,
 * examples/benchmark.jl:18-24, plus CUDA.@profile): it times the kernels of
 * eager steps with HIP events recorded on the simulation's own stream.
 */
#ifndef FCC_BENCH_H
#define FCC_BENCH_H

#include "fcc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Run `steps` eager steps (advancing the simulation like fcc_step) and
 * return the average device time of the forces kernel and of the update
 * kernel(s) of a step, in milliseconds, from hipEventElapsedTime. Steps that
 * rebuild the neighbour list are included; the rebuild itself is not timed.
 * Where a step runs as one kernel (forces and update, see
 * FCC_FLAG_TWO_KERNELS), forces_ms is that kernel's time and update_ms 0. */
int fcc_time_step_kernels(fcc_sim *sim, int64_t steps, double *forces_ms,
                          double *update_ms);

/* fcc_run (graph mode) with HIP events on the simulation's stream; returns
 * average kernel durations in milliseconds. One-kernel steps: events around
 * each chunk's graph launches, forces_ms = the chunk's time over its steps
 * (the step kernel's duration: inside a graph the kernels follow each other
 * without a gap), update_ms 0. Two-kernel steps: the last step of every
 * chunk is launched eagerly (same kernels and arguments as the captured
 * step) with events around each kernel, forces_ms and update_ms the averages
 * over those steps. Chunks with frozen steps are left out. Results equal
 * fcc_run's. */
int fcc_run_timed(fcc_sim *sim, int64_t n, double *forces_ms, double *update_ms);

/* Average duration of the forces kernel and of the update kernel(s) of a
 * step, in milliseconds, from two series between HIP events on the
 * simulation's stream, each a captured graph as in fcc_run (the forces
 * series is first replayed untimed, about 400 launches in all, for the clock
 * ramp of an idle GPU):
 * `reps` back-to-back forces launches (evaluating the current state), then
 * `reps` steps (forces + update, or one-kernel steps where they apply);
 * update = (steps - forces) / reps. The state is saved before and restored
 * afterwards: the simulation is left unchanged. */
int fcc_time_kernels(fcc_sim *sim, int64_t reps, double *forces_ms, double *update_ms);

/* `reps` forced rebuilds of the cell and neighbour lists of the current state
 * (as fcc_update_neighbors(sim, 1)), with HIP events between the stages;
 * stage_ms[5] receives the average duration in milliseconds of: cell index
 * (count clear + cell ids and counts), scan (two launches), scatter, cell
 * sort (with the cell-sorted copies of positions and bounds), neighbour list
 * (with the Verlet reference). The events of a stage are recorded by the
 * dispatches of its first and last kernel (start and end), so the figures are
 * kernel durations (plus the gap between the two launches of a two-kernel
 * stage). Counts as `reps` list builds. */
int fcc_time_rebuild(fcc_sim *sim, int64_t reps, double *stage_ms);

/* fcc_ensemble_run with the same eager last step per chunk between HIP
 * events; average forces and update kernel durations of an ensemble step
 * (all replicas), in milliseconds. Results equal fcc_ensemble_run's. */
int fcc_ensemble_run_timed(fcc_ensemble *ens, int64_t n, double *forces_ms,
                           double *update_ms);

#ifdef __cplusplus
}
#endif

#endif /* FCC_BENCH_H */
