/*
 * fcc_oracle.h — FP64 CPU restatement of the FourierContourCells.jl time step.
 *
 * TEST INFRASTRUCTURE ONLY. This is the parity oracle and the CPU baseline:
 * only tests/, the smoke check of __graft_entry__.py and the cpu_baseline leg
 * of bench.py may load it. The product path (libfcc_hip.so) never links or
 * calls it.
 *
 * It follows the reference's CPU backend (KernelAbstractions CPU(),
 * BodyParallel layout, three BLAS products) function by function; each
 * function in fcc_oracle.c cites the reference file:line it restates. The
 * structs are those of include/fcc.h, so one configuration drives both the
 * oracle and the device library.
 */
#ifndef FCC_ORACLE_H
#define FCC_ORACLE_H

#include "../include/fcc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fcco_sim fcco_sim;

const char *fcco_last_error(void);
int fcco_create(const fcc_config *cfg, const fcc_bodies *bodies, fcco_sim **out);
int fcco_destroy(fcco_sim *s);
int fcco_set_bodies(fcco_sim *s, const fcc_bodies *b);
int fcco_get_bodies(fcco_sim *s, fcc_bodies *b);
int fcco_set_dt(fcco_sim *s, double dt);
int fcco_set_bonds(fcco_sim *s, int64_t n, const int32_t *pairs);
int fcco_step(fcco_sim *s, int64_t n);
int fcco_update_neighbors(fcco_sim *s, int32_t force);
int fcco_compute_gradients(fcco_sim *s, int32_t frame);
int fcco_get_gradients(fcco_sim *s, double *gx, double *gy, double *gth,
                       double *glab, double *gcoef, double *Ebody);
int fcco_get_contour_gradients(fcco_sim *s, double *P, double *GR, double *GD,
                               double *GDD);
int fcco_get_geometry(fcco_sim *s, double *lab, double *rmax, double *R,
                      double *dR, double *ddR);
int fcco_shape_index(fcco_sim *s, double *perimeters, double *areas,
                     double *shape_index);
int fcco_contours(fcco_sim *s, double *X, double *Y);
int fcco_get_celllist(fcco_sim *s, int32_t *ncx, int32_t *ncy, int32_t *cellid,
                      int32_t *perm, int32_t *start, int32_t *stop);
int fcco_get_neighbors(fcco_sim *s, int32_t *capacity, int32_t *count,
                       int32_t *idx);
int fcco_set_neighbor_capacity(fcco_sim *s, int32_t capacity);
int fcco_get_counters(fcco_sim *s, int64_t *step, double *t, int64_t *clock,
                      int64_t *builds);
/* Overwrite the gradients gx, gy (N each) before fcco_integrate_noise. */
int fcco_set_center_gradients(fcco_sim *s, const double *gx, const double *gy);
/* integrate!(bodies, grad, disp, ξ, ...) with a 3 x N noise matrix and
 * frame = false (src/dynamics.jl:220-232): used by the parent-drag tests. */
int fcco_integrate_noise(fcco_sim *s, const double *xi);
/* Build the cell and neighbour lists from given positions and bounds (no
 * simulation state): the bit-exactness check of the device lists. */
int fcco_build_lists(int64_t n, const double *x, const double *y,
                     const double *rmax, double Lx, double Ly, double cutoff,
                     double range, int32_t capacity, int32_t *ncx, int32_t *ncy,
                     int32_t *cellid, int32_t *perm, int32_t *start,
                     int32_t *stop, int32_t *count, int32_t *idx,
                     int32_t *max_count);

/* Random numbers (src/random.jl). */
void fcco_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double fcco_uniform_open64(uint32_t a, uint32_t b);
float fcco_uniform_open32(uint32_t a);
int fcco_philox_noise(uint64_t seed, uint64_t draw, int64_t n, int32_t precision,
                      void *out);

/* Box helpers (src/box.jl:14-30). */
double fcco_minimum_image(double d, double L);
double fcco_wrap(double x, double L, int32_t *shift);

/* Threads used by the OpenMP loops (1 = serial). */
void fcco_set_threads(int n);
int fcco_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif
