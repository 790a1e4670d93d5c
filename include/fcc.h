/*
 * fcc.h — C ABI of libfcc_hip.so, the MI355X (gfx950) time step of the
 * Fourier contour cell model (FourierContourCells.jl).
 *
 * Conventions follow the reference's Julia arrays so that a Julia package
 * extension can pass its `Bodies` fields unchanged:
 *   - all sizes are int64_t, body counts N, mode count K = maxmode - 1,
 *     ncoef = 2K + 1, contour nodes n_theta;
 *   - matrices are column-major with one column per body (coef is ncoef x N,
 *     the neighbour list is capacity x N);
 *   - body types and parent ids are 1-based, as in `Bodies.type` /
 *     `Bodies.parent` (src/bodies.jl:53-72);
 *   - host arrays are owned by the caller and only copied in or out; the
 *     library owns every device buffer of a simulation (src/simulation.jl:91-101).
 * Every entry point returns FCC_OK (0) or a negative error code, with a
 * message in fcc_last_error() (thread-local). Nothing here takes torch or
 * C++ types.
 */
#ifndef FCC_H
#define FCC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FCC_ABI_VERSION 4

#define FCC_MAX_TYPES 4        /* body types per simulation */
#define FCC_MAX_TERMS 8        /* energy terms per simulation */
#define FCC_MAX_MODES 31       /* K = maxmode - 1 */
#define FCC_MAX_THETA 8000     /* contour nodes handled on the device (above 64: one workgroup per body) */

/* Error codes (the reference throws ArgumentError / DimensionMismatch on the
 * host, src/simulation.jl:71-118, src/basis.jl:36-38, src/bodies.jl:76-77). */
enum fcc_status {
    FCC_OK = 0,
    FCC_ERR_ARGUMENT = -1,      /* ArgumentError in the reference */
    FCC_ERR_DIMENSION = -2,     /* DimensionMismatch in the reference */
    FCC_ERR_UNSUPPORTED = -3,   /* valid for the reference, not native here:
                                   the caller falls back to the reference path */
    FCC_ERR_DEVICE = -4,        /* HIP runtime error */
    FCC_ERR_STATE = -5          /* call not valid in the current state */
};

/* Arithmetic of the time step. */
enum fcc_precision {
    FCC_FP64 = 0,        /* everything in FP64 (parity mode, T = Float64) */
    FCC_FP32 = 1,        /* plain FP32 (T = Float32); positions stored in FP64 */
    FCC_MIXED_FP16 = 2,  /* contour contractions on MFMA in FP16 hi+lo with FP32
                            accumulation, contact FP32, a0 and wrap FP64 */
    FCC_MIXED_BF16 = 3   /* as FCC_MIXED_FP16 with BF16 hi+lo */
};

/* Built-in energy terms (src/energies/body_terms.jl, src/energies/contact.jl;
 * Bending as defined in test/test_energies.jl:6-10). */
enum fcc_term_kind {
    FCC_TERM_TENSION = 1,               /* p[0] = eta */
    FCC_TERM_PERIMETER_ELASTICITY = 2,  /* p[0] = lambda, p[1] = p0 */
    FCC_TERM_BENDING = 3,               /* p[0] = kappa */
    FCC_TERM_OVERLAP = 16,              /* J, Jrel tables */
    FCC_TERM_LINEAR_REPULSION = 17,     /* p[0] = k */
    FCC_TERM_HARMONIC_BOND = 32         /* p[0] = k, p[1] = r0 (src/energies/bonds.jl:12-29);
                                           acts on the pairs of fcc_set_bonds */
};

/* Simulation flags. */
enum fcc_flags {
    FCC_FLAG_DEBUG_BUFFERS = 1, /* keep P, GR, GD, GDD and gcoef for inspection */
    FCC_FLAG_TWO_LANES = 2,     /* forces kernels with two lanes per body at every N
                                   (by default four lanes below 131072 bodies at
                                   n_theta = 40); same results up to rounding, for
                                   tests of the two-lane kernels on small systems */
    FCC_FLAG_TWO_KERNELS = 4    /* steps as a forces and an update kernel where the
                                   one-kernel step would apply (single FP32 / mixed
                                   simulations on the two-lane kernels without parent
                                   drag, bonds or debug buffers); bitwise the same
                                   results, for tests and kernel profiles */
};

/* BodyType (src/bodies.jl:30-45). */
typedef struct fcc_body_type {
    double v0, Dr, Dt, mur, mutheta, muab;
    int32_t parent_drag;   /* drag = :parent */
    int32_t reserved;
} fcc_body_type;

/* One energy term; `n_on` > 0 restricts it to the listed (1-based) body
 * types like OnTypes (src/energies/interface.jl:278-284). Overlap couplings
 * are indexed J[(ti-1)*FCC_MAX_TYPES + (tj-1)] (src/energies/contact.jl:12-28). */
typedef struct fcc_term {
    int32_t kind;
    int32_t n_on;
    int32_t on_types[FCC_MAX_TYPES];
    double p[2];
    double J[FCC_MAX_TYPES * FCC_MAX_TYPES];
    double Jrel[FCC_MAX_TYPES * FCC_MAX_TYPES];
} fcc_term;

/* Arguments of Simulation(bodies, box; types, energies, dt, nθ, T, seed, skin)
 * (src/simulation.jl:66-68). */
typedef struct fcc_config {
    int64_t n_bodies;
    int32_t n_modes;       /* K */
    int32_t n_theta;
    int32_t n_types;
    int32_t n_terms;
    fcc_body_type types[FCC_MAX_TYPES];
    fcc_term terms[FCC_MAX_TERMS];
    double dt, Lx, Ly, skin;
    uint64_t seed;
    int32_t precision;     /* enum fcc_precision */
    int32_t flags;         /* enum fcc_flags */
} fcc_config;

/* Host view of `Bodies` (src/bodies.jl:62-72). Any pointer may be NULL in
 * fcc_get_bodies to skip that field. */
typedef struct fcc_bodies {
    double *x, *y, *theta;
    int32_t *imgx, *imgy;
    double *coef;          /* ncoef x N */
    double *R0;
    int32_t *type;         /* 1-based */
    int32_t *parent;       /* 1-based */
} fcc_bodies;

typedef struct fcc_sim fcc_sim;

int fcc_abi_version(void);
const char *fcc_last_error(void);

/* Simulation(...) (src/simulation.jl:66-105): validates like the reference
 * (check_parents :108-118), enforces the shape constraint and computes the
 * lab-frame geometry. FCC_ERR_UNSUPPORTED for terms or sizes without a
 * native kernel. */
int fcc_create(const fcc_config *cfg, const fcc_bodies *bodies, fcc_sim **out);
int fcc_destroy(fcc_sim *sim);

/* Replace the body state, then re-enforce the shape, recompute the geometry
 * and invalidate the neighbour list (the test helper refresh!,
 * test/runtests.jl:9-15, minus the gradient evaluation). */
int fcc_set_bodies(fcc_sim *sim, const fcc_bodies *bodies);
/* The time step stores the shape as lab-frame rows; the body-frame `coef`
 * returned here is derived from them (rotation by -nθ, within rounding of
 * the coefficients the step computed). */
int fcc_get_bodies(fcc_sim *sim, fcc_bodies *bodies);
int fcc_set_dt(fcc_sim *sim, double dt);

/* The `bonds` argument of Simulation (src/simulation.jl:14-15,96): n pairs
 * (i, j) of 1-based body ids, pairs[2q], pairs[2q + 1], on which the bond
 * terms act; built into the per-body partner list of BondList
 * (src/energies/bonds.jl:47-76) with its checks (FCC_ERR_ARGUMENT for a body
 * outside 1:N, a self bond or a duplicate). n = 0 removes all bonds. */
int fcc_set_bonds(fcc_sim *sim, int64_t n, const int32_t *pairs);

/* step!(sim, n) (src/simulation.jl:165-173): eager steps, one host
 * synchronisation per step for the skin check. */
int fcc_step(fcc_sim *sim, int64_t n);

/* run_steps!(backend, sim, n) in graph mode (ext/FourierContourCellsCUDAExt.jl:144-180):
 * the step is captured in a hipGraph and replayed in chunks with one host
 * synchronisation per chunk; results are bitwise those of fcc_step. */
int fcc_run(fcc_sim *sim, int64_t n);

/* update_neighbors!(sim; force) (src/simulation.jl:132-142). */
int fcc_update_neighbors(fcc_sim *sim, int32_t force);

/* compute_gradients!(sim; frame) (src/simulation.jl:150-158). */
int fcc_compute_gradients(fcc_sim *sim, int32_t frame);

/* Gradients (src/gradients.jl:15-27): gx, gy, gth, Ebody [N]; glab, gcoef
 * [ncoef x N]. NULL pointers are skipped. gcoef needs frame = 1 in the last
 * fcc_compute_gradients or FCC_FLAG_DEBUG_BUFFERS during steps. */
int fcc_get_gradients(fcc_sim *sim, double *gx, double *gy, double *gth,
                      double *glab, double *gcoef, double *Ebody);

/* Contour gradients and contact density [n_theta x N] (needs
 * FCC_FLAG_DEBUG_BUFFERS). */
int fcc_get_contour_gradients(fcc_sim *sim, double *P, double *GR, double *GD,
                              double *GDD);

/* Geometry (src/geometry.jl:13-32): lab [ncoef x N], rmax [N], and the
 * contours R, dR, ddR [n_theta x N] synthesised on demand. */
int fcc_get_geometry(fcc_sim *sim, double *lab, double *rmax, double *R,
                     double *dR, double *ddR);

/* perimeters / areas / shape_index of all bodies (src/observables.jl:6-18). */
int fcc_shape_index(fcc_sim *sim, double *perimeters, double *areas,
                    double *shape_index);

/* Cell list (src/neighbors.jl:16-33) in the reference's 1-based convention:
 * cellid, perm [N]; start, stop [ncx*ncy] (empty cells 1/0). */
int fcc_get_celllist(fcc_sim *sim, int32_t *ncx, int32_t *ncy, int32_t *cellid,
                     int32_t *perm, int32_t *start, int32_t *stop);

/* Neighbour list (src/neighbors.jl:110-127): capacity and count [N]; idx
 * [capacity x N] (1-based body ids) if not NULL. */
int fcc_get_neighbors(fcc_sim *sim, int32_t *capacity, int32_t *count,
                      int32_t *idx);

/* Shrink or grow the neighbour list matrix and invalidate it, like assigning
 * sim.neighbors.idx in test/test_contact.jl:81 and test/test_cuda.jl:72. */
int fcc_set_neighbor_capacity(fcc_sim *sim, int32_t capacity);

/* Counters: step, t (host), device clock (Philox draw index), builds. */
int fcc_get_counters(fcc_sim *sim, int64_t *step, double *t, int64_t *clock,
                     int64_t *builds);

/* philox_noise!(ξ, philox_key(seed), draw) (src/random.jl:113-130): 3 x n
 * standard normals, float (FCC_FP32) or double (FCC_FP64), computed on the
 * device into host memory `out`. */
int fcc_philox_noise(uint64_t seed, uint64_t draw, int64_t n, int32_t precision,
                     void *out);

/* fcc_philox_noise without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc). The device staging buffer is a
 * fixed array of the library, one per GPU (3 x FCC_NOALLOC_MAX_NOISE doubles,
 * 6 MiB, held until the process ends); n above FCC_NOALLOC_MAX_NOISE returns
 * FCC_ERR_ARGUMENT. Runs on the current device's null stream and returns when
 * `out` is filled; every call on one GPU shares its buffer, so calls that
 * overlap on one GPU (from several threads) race. */
#define FCC_NOALLOC_MAX_NOISE 262144
int fcc_philox_noise_noalloc(uint64_t seed, uint64_t draw, int64_t n, int32_t precision,
                             void *out);

/* Raw Philox4x32-10 blocks on the device (src/random.jl:25-36): ctr [4 x n],
 * key [2 x n] -> out [4 x n]. */
int fcc_philox4x32(const uint32_t *ctr, const uint32_t *key, int64_t n,
                   uint32_t *out);

/* contours(sim) (src/observables.jl:36-44): contour node coordinates around
 * the wrapped centres, X, Y [n_theta x N] (synchronises). */
int fcc_contours(fcc_sim *sim, double *X, double *Y);

/* Frames for live views (the Makie producer of
 * ext/FourierContourCellsMakieExt.jl:209-276 pulls geometry after every
 * run!): fcc_frame_request enqueues the contours of the current state on the
 * simulation's stream and their copy to pinned host memory on a copy stream,
 * and returns; steps issued afterwards run while the copy is in flight.
 * fcc_frame_read waits for that frame and copies it out (X, Y as
 * fcc_contours, clock = the device clock of the frame). One frame is held at
 * a time; FCC_ERR_STATE if none was requested. */
int fcc_frame_request(fcc_sim *sim);
int fcc_frame_read(fcc_sim *sim, double *X, double *Y, int64_t *clock);

/* ---- checkpoint and resume
 *
 * The reference has no checkpoint; a resumed Simulation built from saved
 * bodies restarts its device clock (the Philox draw index,
 * src/simulation.jl:38-64,99) and its Verlet reference
 * (src/neighbors.jl:110-119), so the continued run differs from the
 * uninterrupted one. A checkpoint here holds the body state, the lab frame,
 * the Verlet reference, the neighbour list (whose order fixes the contact
 * summation order), the device clock, the rebuild flag and the step, time
 * and build counters: loading it into a simulation created with the same
 * configuration, body types and parents continues bitwise as the saved one
 * would have. The blob is this library's format (native byte order). */

/* Write the checkpoint into buf (synchronises). *needed receives its size;
 * buf = NULL only queries it. FCC_ERR_DIMENSION if bytes is too small. */
int fcc_checkpoint_save(fcc_sim *sim, void *buf, int64_t bytes, int64_t *needed);

/* Restore a checkpoint. FCC_ERR_DIMENSION for another precision, size or
 * truncated data; FCC_ERR_ARGUMENT for another dt, box, skin, seed, other
 * body types or parents, or data that is not a checkpoint. */
int fcc_checkpoint_load(fcc_sim *sim, const void *buf, int64_t bytes);

/* ---- long-run observables on the device
 *
 * The reference evaluates shape_index and unwrapped_positions on the host
 * from full copies of the state (src/observables.jl:6-29), typically in a
 * run!(sim, n; every, callback) loop (src/simulation.jl:220-235). Here a
 * sample is enqueued on the simulation's stream without a host transfer and
 * accumulates, per body type t (1-based t = index + 1):
 *   - the mean squared displacement of the unwrapped centres from their
 *     positions at fcc_observer_start, mean over the bodies of type t;
 *   - the mean shape index p = perimeter / sqrt(area) and the mean of p^2;
 *   - a histogram of p pooled over all samples: bin 0 holds p < pmin, bin
 *     b in 1..nbins holds 1 + floor((p - pmin) * nbins / (pmax - pmin)) == b,
 *     bin nbins + 1 holds p >= pmax (and NaN).
 * p is the value fcc_shape_index returns, computed in the simulation's
 * precision; sums are FP64 and reduced in a fixed order. */
#define FCC_OBS_MAX_BINS 2046  /* FCC_MAX_TYPES x (nbins + 2) counters in 32 KiB of LDS */

typedef struct fcc_observer_config {
    int32_t nbins;         /* shape-index bins, 1..FCC_OBS_MAX_BINS */
    int32_t max_samples;   /* length of the sample series, >= 1 */
    double pmin, pmax;     /* histogram range, pmin < pmax */
} fcc_observer_config;

/* Start (or restart) an observation: the current unwrapped positions become
 * the origin of the displacements; the series and the histogram are cleared. */
int fcc_observer_start(fcc_sim *sim, const fcc_observer_config *cfg);

/* Take one sample of the current state (asynchronous). FCC_ERR_STATE without
 * an observation or when max_samples samples have been taken. */
int fcc_observe(fcc_sim *sim);

/* run!(sim, n; every, callback) with a device sample as the callback: n
 * steps in runs of `every` steps, one sample after each. */
int fcc_run_observed(fcc_sim *sim, int64_t n, int64_t every);

/* Copy the observation out (synchronises): n_samples; per sample the device
 * clock (steps since creation), and msd, p_mean, p2_mean as [n_types x
 * n_samples] column-major (type fastest); hist as [(nbins + 2) x n_types]
 * column-major (bin fastest). Arrays are sized for max_samples; any pointer
 * may be NULL. A type without bodies reads NaN. */
int fcc_observer_read(fcc_sim *sim, int32_t *n_samples, int64_t *clock, double *msd,
                      double *p_mean, double *p2_mean, uint64_t *hist);

/* ---- ensembles of independent simulations
 *
 * Independent simulations advanced together, in place of one task and stream
 * per simulation (README.md:249-261, test/test_cuda.jl:95-104). The replicas
 * share precision, n_modes, n_theta, parent drag (present or not) and flags;
 * N, box, body types, energy parameters, dt, skin and seed are their own.
 * Each step of all replicas is one launch per kernel over a table of
 * (replica, workgroup) tiles, captured once in a hipGraph; a replica whose
 * Verlet skin trips is frozen until its list is rebuilt between graph
 * chunks. Every replica ends bitwise where fcc_run would take it alone. */
typedef struct fcc_ensemble fcc_ensemble;

/* Launch shape of the ensemble step (set by hand or by the autotuner). */
typedef struct {
    int32_t update_block; /* bodies per workgroup of the update kernel: 64, 128 or 256 */
    int32_t xcd_order;    /* 1: each XCD runs one contiguous range of tiles (8 XCDs) */
    int32_t chunk;        /* graph replays between host status reads; 0 = adaptive */
    int32_t lanes;        /* lanes per body of the n_theta = 40 forces kernels: 0 = each
                             replica's own (four below 131072 bodies, two with
                             FCC_FLAG_TWO_LANES), 2 = two for every replica from now on, as
                             if created with FCC_FLAG_TWO_LANES (member handles included) */
    int32_t kernels;      /* 0: one kernel per step where every replica qualifies (two lanes,
                             FP32 or mixed without Bending, no parent drag, bonds or debug
                             buffers; see FCC_FLAG_TWO_KERNELS), 2: a forces and an update
                             kernel per step; results are the same bit for bit */
} fcc_launch;

/* Create n replicas from cfgs[n] and bodies[n] (validated like fcc_create).
 * FCC_ERR_UNSUPPORTED if the replicas differ in what must be shared. */
int fcc_ensemble_create(int32_t n, const fcc_config *cfgs, const fcc_bodies *bodies,
                        fcc_ensemble **out);
int fcc_ensemble_destroy(fcc_ensemble *ens);

/* Borrowed handle of replica r (0-based), valid until fcc_ensemble_destroy:
 * every fcc_* call applies to it except fcc_destroy. */
int fcc_ensemble_member(fcc_ensemble *ens, int32_t r, fcc_sim **member);

/* Advance every replica by n steps (run!(sim, n) on each). */
int fcc_ensemble_run(fcc_ensemble *ens, int64_t n);

int fcc_ensemble_get_launch(fcc_ensemble *ens, fcc_launch *launch);
int fcc_ensemble_set_launch(fcc_ensemble *ens, const fcc_launch *launch);

/* Autotuner: replays `steps` ensemble steps for every candidate launch shape
 * (update_block x xcd_order, lanes 0 / 2 when some replica would take the
 * four-lane kernels, kernels 0 / 2 where one kernel per step applies) from
 * the same state, restored after each trial and at the end (the run state is
 * unchanged), keeps the fastest and returns it with its device time per step
 * in milliseconds. With lanes = 2 chosen, the replicas continue as replicas
 * created with FCC_FLAG_TWO_LANES would. */
int fcc_ensemble_autotune(fcc_ensemble *ens, int64_t steps, fcc_launch *best,
                          double *ms_per_step);

/* Observations of every replica (fcc_observer_start / fcc_observe on each
 * member; fcc_observer_read per member handle). */
int fcc_ensemble_observer_start(fcc_ensemble *ens, const fcc_observer_config *cfg);
int fcc_ensemble_observe(fcc_ensemble *ens);

/* fcc_ensemble_run in runs of `every` steps, every replica sampled after each. */
int fcc_ensemble_run_observed(fcc_ensemble *ens, int64_t n, int64_t every);

#ifdef __cplusplus
}
#endif

#endif /* FCC_H */
