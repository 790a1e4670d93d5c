/*
 * fcc_oracle.c — FP64 CPU restatement of the FourierContourCells.jl time step.
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle and CPU baseline; see fcc_oracle.h).
 *
 * Every routine restates the reference's CPU backend (BodyParallel kernels,
 * three BLAS products) in plain C with IEEE double arithmetic and no FMA
 * contraction (built with -ffp-contract=off, as Julia does not contract),
 * citing the reference file:line it follows. Arrays are column-major with
 * one column per body, like the Julia arrays. Internally body, type, cell and
 * neighbour indices are 0-based; the accessors convert to the reference's
 * 1-based values.
 */
#include "fcc_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define STATUS_REBUILD 0
#define STATUS_UPDATE 1

static _Thread_local char g_err[512];
static int g_threads = 1;

static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *fcco_last_error(void) { return g_err; }
void fcco_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int fcco_get_threads(void) { return g_threads; }

struct fcco_sim {
    int64_t N;
    int K, nc, nth, ntypes;
    double dphi;
    double *C, *D, *DD, *cs;         /* basis: nth x nc, cs 2 x nth */
    double Lx, Ly;
    /* bodies */
    double *x, *y, *th, *coef, *R0;
    int32_t *imgx, *imgy, *type, *parent;
    /* geometry */
    double *lab, *labR0, *R, *dR, *ddR, *rmax;
    /* gradients */
    double *GR, *GD, *GDD, *P, *gx, *gy, *gth, *glab, *gcoef, *Ebody;
    /* cell list */
    int32_t ncx, ncy;
    double wx, wy;
    int32_t *cellid, *perm, *start, *stop, *scratch;
    /* neighbour list */
    int32_t cap;
    int32_t *nidx, *ncount;
    double skin;
    double *x0, *y0, *rmax0;
    int valid;
    int64_t builds;
    /* children */
    int32_t ccap;
    int32_t *cidx, *ccount;
    /* terms and types */
    int nbody, ncontact, nbondt;
    fcc_term body[FCC_MAX_TERMS], contact[FCC_MAX_TERMS], bondt[FCC_MAX_TERMS];
    /* bond list (BondList, src/energies/bonds.jl:47-76): partners of body i in
     * bidx[q + bcap * i], q < bcount[i] */
    int32_t bcap;
    int32_t *bidx, *bcount;
    double range;
    fcc_body_type types[FCC_MAX_TYPES];
    uint32_t key[2];
    double *disp;
    int32_t status[3];
    int64_t clock;
    double dt, t;
    int64_t step;
};

/* ---------------------------------------------------------------- random */

/* Philox4x32-10, src/random.jl:10-36 (Salmon et al., SC'11). */
void fcco_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* uniform_open, src/random.jl:42-44. */
double fcco_uniform_open64(uint32_t a, uint32_t b) {
    uint64_t j = (((uint64_t)a << 32) | (uint64_t)b) >> 12;
    return ((double)j + 0.5) * 0x1p-52;
}
float fcco_uniform_open32(uint32_t a) { return ((float)(a >> 9) + 0.5f) * 0x1p-23f; }

/* box_muller + philox_pair, src/random.jl:47-51,75-78. */
static void philox_pair64(const uint32_t key[2], int64_t i1, uint64_t n, uint32_t block,
                          double *z1, double *z2) {
    uint32_t ctr[4] = {(uint32_t)i1, (uint32_t)n, (uint32_t)(n >> 32), block}, a[4];
    fcco_philox4x32(ctr, key, a);
    double u1 = fcco_uniform_open64(a[0], a[1]);
    double u2 = fcco_uniform_open64(a[2], a[3]);
    double r = sqrt(-2.0 * log(u1));
    double ang = 2.0 * 3.141592653589793 * u2;
    *z1 = r * cos(ang);
    *z2 = r * sin(ang);
}

static void philox_pair32(const uint32_t key[2], int64_t i1, uint64_t n, uint32_t block,
                          float *z1, float *z2) {
    uint32_t ctr[4] = {(uint32_t)i1, (uint32_t)n, (uint32_t)(n >> 32), block}, a[4];
    fcco_philox4x32(ctr, key, a);
    float u1 = fcco_uniform_open32(a[0]);
    float u2 = fcco_uniform_open32(a[2]);
    float r = sqrtf(-2.0f * logf(u1));
    float ang = 2.0f * 3.14159274f * u2;
    *z1 = r * cosf(ang);
    *z2 = r * sinf(ang);
}

/* philox_key, src/random.jl:103. */
static void philox_key(uint64_t seed, uint32_t key[2]) {
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32);
}

/* noise_kernel! / philox_noise!, src/random.jl:113-130: 3 x n, body i uses
 * counter i (1-based). */
int fcco_philox_noise(uint64_t seed, uint64_t draw, int64_t n, int32_t precision, void *out) {
    uint32_t key[2];
    philox_key(seed, key);
    if (precision == FCC_FP64) {
        double *o = (double *)out;
        for (int64_t i = 0; i < n; i++) {
            double z1, z2, z3, z4;
            philox_pair64(key, i + 1, draw, 0, &z1, &z2);
            philox_pair64(key, i + 1, draw, 1, &z3, &z4);
            o[3 * i] = z1;
            o[3 * i + 1] = z2;
            o[3 * i + 2] = z3;
        }
    } else if (precision == FCC_FP32) {
        float *o = (float *)out;
        for (int64_t i = 0; i < n; i++) {
            float z1, z2, z3, z4;
            philox_pair32(key, i + 1, draw, 0, &z1, &z2);
            philox_pair32(key, i + 1, draw, 1, &z3, &z4);
            o[3 * i] = z1;
            o[3 * i + 1] = z2;
            o[3 * i + 2] = z3;
        }
    } else {
        return fail(FCC_ERR_ARGUMENT, "precision must be FP64 or FP32");
    }
    return FCC_OK;
}

/* ------------------------------------------------------------------- box */

/* minimum_image, src/box.jl:14 (round is ties-to-even: rint). */
double fcco_minimum_image(double d, double L) { return d - L * rint(d / L); }

/* wrap, src/box.jl:21-30. */
double fcco_wrap(double x, double L, int32_t *shift) {
    double s = floor(x / L);
    double xp = x - s * L;
    if (xp >= L) {
        xp = 0.0;
        s += 1.0;
    }
    *shift = (int32_t)s;
    return xp;
}

/* ----------------------------------------------------------------- basis */

/* FourierBasis{T}, src/basis.jl:35-63. */
static void basis_init(fcco_sim *s) {
    int nth = s->nth, K = s->K, nc = s->nc;
    double dphi = 2.0 * 3.141592653589793 / (double)nth;
    s->dphi = dphi;
    for (int k = 0; k < nth; k++) {
        double phi = (double)k * dphi;
        s->cs[2 * k] = cos(phi);
        s->cs[2 * k + 1] = sin(phi);
        s->C[k] = 1.0;
        s->D[k] = 0.0;
        s->DD[k] = 0.0;
        for (int m = 1; m <= K; m++) {
            int n = m + 1;
            double sn = sin((double)n * phi), cn = cos((double)n * phi);
            s->C[k + nth * m] = cn;
            s->C[k + nth * (K + m)] = sn;
            s->D[k + nth * m] = (double)(-n) * sn;
            s->D[k + nth * (K + m)] = (double)n * cn;
            s->DD[k + nth * m] = (double)(-n * n) * cn;
            s->DD[k + nth * (K + m)] = (double)(-n * n) * sn;
        }
    }
    (void)nc;
}

/* -------------------------------------------------------------- geometry */

/* lab_frame!, src/geometry.jl:73-92. */
static void lab_frame(fcco_sim *s, int64_t i, double thi) {
    int K = s->K, nc = s->nc;
    double r0 = s->R0[i];
    const double *c = s->coef + nc * i;
    double *lab = s->lab + nc * i, *labR0 = s->labR0 + nc * i;
    lab[0] = c[0];
    labR0[0] = r0 * c[0];
    double bound = c[0];
    for (int m = 1; m <= K; m++) {
        int n = m + 1;
        double a = c[m], b = c[K + m];
        double sn = sin((double)n * thi), cn = cos((double)n * thi);
        double al = a * cn - b * sn;
        double be = a * sn + b * cn;
        lab[m] = al;
        lab[K + m] = be;
        labR0[m] = r0 * al;
        labR0[K + m] = r0 * be;
        bound += sqrt(a * a + b * b);
    }
    s->rmax[i] = r0 * bound;
}

/* synthesize! on the CPU (three products), src/geometry.jl:113-122. */
static void synthesize(fcco_sim *s) {
    int nth = s->nth, nc = s->nc;
    int64_t N = s->N;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t i = 0; i < N; i++) {
        const double *l = s->labR0 + nc * i;
        for (int k = 0; k < nth; k++) {
            double r = 0.0, dr = 0.0, ddr = 0.0;
            for (int c = 0; c < nc; c++) {
                r += s->C[k + nth * c] * l[c];
                dr += s->D[k + nth * c] * l[c];
                ddr += s->DD[k + nth * c] * l[c];
            }
            s->R[k + nth * i] = r;
            s->dR[k + nth * i] = dr;
            s->ddR[k + nth * i] = ddr;
        }
    }
}

/* update_geometry!, src/geometry.jl:105-110. */
static void update_geometry(fcco_sim *s) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t i = 0; i < s->N; i++) lab_frame(s, i, s->th[i]);
    synthesize(s);
}

/* contour_radius_dir, src/geometry.jl:147-156. */
static double contour_radius_dir(const double *labj, int K, double R0, double c1, double s1) {
    double c = c1, sn = s1, r = labj[0];
    for (int m = 1; m <= K; m++) {
        double cn = c * c1 - sn * s1;
        double snn = sn * c1 + c * s1;
        c = cn;
        sn = snn;
        r += labj[m] * c + labj[K + m] * sn;
    }
    return R0 * r;
}

/* --------------------------------------------------------------- dynamics */

#define SHAPE_BOUND 0.816496580927726 /* sqrt(2/3), src/dynamics.jl:26 */

/* enforce_shape!, src/dynamics.jl:44-64. */
static void enforce_shape(double *c, int K) {
    double sm = 0.0;
    for (int m = 1; m <= K; m++) sm += sqrt(c[m] * c[m] + c[K + m] * c[K + m]);
    double bound = sqrt(2.0 / 3.0);
    if (sm > bound) {
        double f = bound / sm;
        for (int m = 1; m <= K; m++) {
            c[m] *= f;
            c[K + m] *= f;
        }
    }
    double q = 0.0;
    for (int m = 1; m <= K; m++) q += c[m] * c[m] + c[K + m] * c[K + m];
    c[0] = sqrt(1.0 - q / 2.0);
}

/* ------------------------------------------------------------ body terms */

static int applies(const fcc_term *t, int32_t type0) {
    if (t->n_on <= 0) return 1;
    for (int q = 0; q < t->n_on && q < FCC_MAX_TYPES; q++)
        if (t->on_types[q] == type0 + 1) return 1;
    return 0;
}

/* contour_integrands of the built-in body terms (body_terms.jl:12,33;
 * Bending: test/test_energies.jl:9). */
static double integrand(const fcc_term *t, double R, double dR, double ddR) {
    switch (t->kind) {
    case FCC_TERM_TENSION:
    case FCC_TERM_PERIMETER_ELASTICITY:
        return sqrt(R * R + dR * dR);
    case FCC_TERM_BENDING: {
        double s2 = R * R + dR * dR;
        double num = R * R + 2.0 * dR * dR - R * ddR;
        return num * num / pow(s2, 2.5);
    }
    }
    return 0.0;
}

/* contour_integrand_gradients (Tension body_terms.jl:14-17; the others by
 * their closed forms, which ForwardDiff reproduces to rounding). */
static void integrand_grad(const fcc_term *t, double R, double dR, double ddR, double J[3]) {
    switch (t->kind) {
    case FCC_TERM_TENSION: {
        double s = sqrt(R * R + dR * dR);
        J[0] = R / s;
        J[1] = dR / s;
        J[2] = 0.0;
        return;
    }
    case FCC_TERM_PERIMETER_ELASTICITY: {
        double s = sqrt(R * R + dR * dR);
        double inv = 1.0 / (2.0 * s);
        J[0] = (2.0 * R) * inv;
        J[1] = (2.0 * dR) * inv;
        J[2] = 0.0;
        return;
    }
    case FCC_TERM_BENDING: {
        double s2 = R * R + dR * dR;
        double num = R * R + 2.0 * dR * dR - R * ddR;
        double s5 = pow(s2, -2.5), s7 = pow(s2, -3.5);
        J[0] = 2.0 * num * (2.0 * R - ddR) * s5 - 5.0 * num * num * R * s7;
        J[1] = 2.0 * num * (4.0 * dR) * s5 - 5.0 * num * num * dR * s7;
        J[2] = -2.0 * num * R * s5;
        return;
    }
    }
    J[0] = J[1] = J[2] = 0.0;
}

/* energy and energy_gradient (body_terms.jl:13,18,34; interface.jl:192). */
static double term_energy(const fcc_term *t, double Q, double R0, double *dF) {
    switch (t->kind) {
    case FCC_TERM_TENSION:
        *dF = t->p[0];
        return t->p[0] * Q;
    case FCC_TERM_PERIMETER_ELASTICITY: {
        double x = Q - t->p[1] * sqrt(3.141592653589793) * R0;
        *dF = t->p[0] * (2.0 * x);
        return t->p[0] * (x * x);
    }
    case FCC_TERM_BENDING:
        *dF = t->p[0];
        return t->p[0] * Q;
    }
    *dF = 0.0;
    return 0.0;
}

/* _body_term!, src/gradients.jl:82-98. */
static double body_term(fcco_sim *s, const fcc_term *t, int64_t i) {
    int nth = s->nth;
    if (!applies(t, s->type[i])) return 0.0;
    const double *R = s->R + nth * i, *dR = s->dR + nth * i, *ddR = s->ddR + nth * i;
    double *GR = s->GR + nth * i, *GD = s->GD + nth * i, *GDD = s->GDD + nth * i;
    double Q = integrand(t, R[0], dR[0], ddR[0]);
    for (int k = 1; k < nth; k++) Q += integrand(t, R[k], dR[k], ddR[k]);
    Q *= s->dphi;
    double dF;
    double E = term_energy(t, Q, s->R0[i], &dF);
    double w = dF * s->dphi;
    for (int k = 0; k < nth; k++) {
        double J[3];
        integrand_grad(t, R[k], dR[k], ddR[k], J);
        GR[k] += J[0] * w;
        GD[k] += J[1] * w;
        GDD[k] += J[2] * w;
    }
    return E;
}

/* body_terms_kernel! (BodyParallel) / reset!, src/gradients.jl:47-64,113-128,243-261. */
static void body_terms(fcco_sim *s) {
    int nth = s->nth;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t i = 0; i < s->N; i++) {
        for (int k = 0; k < nth; k++) {
            s->GR[k + nth * i] = 0.0;
            s->GD[k + nth * i] = 0.0;
            s->GDD[k + nth * i] = 0.0;
        }
        s->gx[i] = 0.0;
        s->gy[i] = 0.0;
        s->gth[i] = 0.0;
        double e[FCC_MAX_TERMS];
        for (int q = 0; q < s->nbody; q++) e[q] = body_term(s, &s->body[q], i);
        double E = 0.0; /* right fold e1 + (e2 + (... + 0)) */
        for (int q = s->nbody - 1; q >= 0; q--) E = e[q] + E;
        s->Ebody[i] = E;
    }
}

/* --------------------------------------------------------------- contact */

/* pressure / _pressure, src/energies/contact.jl:25-28,41,46-51. */
static double pressure(const fcco_sim *s, double d, int32_t ti, int32_t tj, int rel) {
    double g[FCC_MAX_TERMS];
    for (int q = 0; q < s->ncontact; q++) {
        const fcc_term *t = &s->contact[q];
        g[q] = 0.0;
        if (!applies(t, ti)) continue;
        if (t->kind == FCC_TERM_OVERLAP) {
            double J = rel ? t->Jrel[ti * FCC_MAX_TYPES + tj] : t->J[ti * FCC_MAX_TYPES + tj];
            g[q] = d >= 0.0 ? J : 0.0;
        } else if (t->kind == FCC_TERM_LINEAR_REPULSION) {
            g[q] = d > 0.0 ? t->p[0] * d : 0.0;
        }
    }
    double acc = 0.0;
    for (int q = s->ncontact - 1; q >= 0; q--) acc = g[q] + acc;
    return acc;
}

/* _node_pressure, src/energies/contact.jl:58-82. */
static double node_pressure(const fcco_sim *s, int k, int64_t i) {
    int nth = s->nth, nc = s->nc;
    double c = s->cs[2 * k], sn = s->cs[2 * k + 1];
    double Rk = s->R[k + nth * i];
    double ux = s->x[i] + Rk * c;
    double uy = s->y[i] + Rk * sn;
    int32_t ti = s->type[i], pi = s->parent[i];
    double p = 0.0;
    int32_t cnt = s->ncount[i];
    for (int32_t q = 0; q < cnt; q++) {
        int32_t j = s->nidx[q + (int64_t)s->cap * i];
        double dx = fcco_minimum_image(ux - s->x[j], s->Lx);
        double dy = fcco_minimum_image(uy - s->y[j], s->Ly);
        double dist2 = dx * dx + dy * dy;
        double lim = s->rmax[j] + s->range;
        if (dist2 > lim * lim) continue;
        double dist = sqrt(dist2);
        double cp = dist > 0.0 ? dx / dist : 1.0;
        double sp = dist > 0.0 ? dy / dist : 0.0;
        double d = contour_radius_dir(s->lab + nc * j, s->K, s->R0[j], cp, sp) - dist;
        p += pressure(s, d, ti, s->type[j], pi == s->parent[j]);
    }
    return p;
}

/* contact_node_kernel!, src/energies/contact.jl:149-187 (fused P, GR and the
 * node-1 serial sums of the centre force and torque). */
static void contact_gradients(fcco_sim *s) {
    int nth = s->nth;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t i = 0; i < s->N; i++) {
        double fx = 0.0, fy = 0.0, tt = 0.0;
        for (int k = 0; k < nth; k++) {
            double g = node_pressure(s, k, i);
            s->P[k + nth * i] = g;
            double p = g * s->dphi;
            double r = s->R[k + nth * i], rp = s->dR[k + nth * i];
            double c = s->cs[2 * k], sn = s->cs[2 * k + 1];
            s->GR[k + nth * i] += p * r;
            fx += p * (r * c + rp * sn);
            fy += p * (r * sn - rp * c);
            tt += (-p) * r * rp;
        }
        s->gx[i] += fx;
        s->gy[i] += fy;
        s->gth[i] += tt;
    }
}

/* ------------------------------------------------------------- pullback */

/* body_frame!, src/gradients.jl:347-362. */
static void body_frame(fcco_sim *s, int64_t i, double thi) {
    int K = s->K, nc = s->nc;
    double r0 = s->R0[i];
    const double *gl = s->glab + nc * i, *c = s->coef + nc * i;
    double *gc = s->gcoef + nc * i;
    double ga0 = r0 * gl[0];
    double a0 = c[0];
    for (int m = 1; m <= K; m++) {
        int n = m + 1;
        double ga = r0 * gl[m], gb = r0 * gl[K + m];
        double sn = sin((double)n * thi), cn = cos((double)n * thi);
        double a = c[m], b = c[K + m];
        gc[m] = ga * cn + gb * sn - ga0 * a / (2.0 * a0);
        gc[K + m] = (-ga) * sn + gb * cn - ga0 * b / (2.0 * a0);
    }
    gc[0] = ga0;
}

/* pullback! on the CPU (three accumulating products), src/gradients.jl:374-386. */
static void pullback(fcco_sim *s, int frame) {
    int nth = s->nth, nc = s->nc;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t i = 0; i < s->N; i++) {
        const double *GR = s->GR + nth * i, *GD = s->GD + nth * i, *GDD = s->GDD + nth * i;
        double *gl = s->glab + nc * i;
        for (int c = 0; c < nc; c++) {
            double a = 0.0;
            for (int k = 0; k < nth; k++) a += s->C[k + nth * c] * GR[k];
            gl[c] = a;
        }
        for (int c = 0; c < nc; c++) {
            double a = 0.0;
            for (int k = 0; k < nth; k++) a += s->D[k + nth * c] * GD[k];
            gl[c] = a + gl[c];
        }
        for (int c = 0; c < nc; c++) {
            double a = 0.0;
            for (int k = 0; k < nth; k++) a += s->DD[k + nth * c] * GDD[k];
            gl[c] = a + gl[c];
        }
        if (frame) body_frame(s, i, s->th[i]);
    }
}

/* ------------------------------------------------------- neighbour search */

static int ncells(double L, double cutoff) {
    int n = (int)floor(L / cutoff);
    return n >= 3 ? n : 1;
}

static int32_t imod(int32_t a, int32_t n) {
    int32_t r = a % n;
    return r < 0 ? r + n : r;
}

/* grid_coords / cell_index_kernel! / sortperm! (stable) / cell_bounds_kernel!,
 * src/neighbors.jl:41-92. The stable sort is a counting sort over cells with
 * bodies visited in ascending order. */
static void celllist_build(int64_t N, const double *x, const double *y, int32_t ncx, int32_t ncy,
                           double wx, double wy, int32_t *cellid, int32_t *perm, int32_t *start,
                           int32_t *stop, int32_t *scratch) {
    int64_t nce = (int64_t)ncx * ncy;
    for (int64_t i = 0; i < N; i++) {
        int32_t cx = (int32_t)(x[i] / wx), cy = (int32_t)(y[i] / wy);
        if (cx > ncx - 1) cx = ncx - 1;
        if (cy > ncy - 1) cy = ncy - 1;
        if (cx < 0) cx = 0;
        if (cy < 0) cy = 0;
        cellid[i] = cx + ncx * cy;
    }
    for (int64_t c = 0; c < nce; c++) scratch[c] = 0;
    for (int64_t i = 0; i < N; i++) scratch[cellid[i]]++;
    int32_t acc = 0;
    for (int64_t c = 0; c < nce; c++) {
        int32_t n = scratch[c];
        scratch[c] = acc;
        acc += n;
    }
    for (int64_t i = 0; i < N; i++) perm[scratch[cellid[i]]++] = (int32_t)i;
    for (int64_t c = 0; c < nce; c++) {
        start[c] = 0;
        stop[c] = -1;
    }
    for (int64_t q = 0; q < N; q++) {
        int32_t c = cellid[perm[q]];
        if (q == 0 || cellid[perm[q - 1]] != c) start[c] = (int32_t)q;
        if (q == N - 1 || cellid[perm[q + 1]] != c) stop[c] = (int32_t)q;
    }
}

/* neighbor_kernel!, src/neighbors.jl:148-174. Returns the largest count. */
static int32_t neighbor_build(int64_t N, const double *x, const double *y, const double *rmax,
                              const int32_t *perm, const int32_t *start, const int32_t *stop,
                              double Lx, double Ly, int32_t ncx, int32_t ncy, double wx, double wy,
                              double range, int32_t cap, int32_t *idx, int32_t *count) {
    int32_t nmax = 0;
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(max : nmax)
    for (int64_t i = 0; i < N; i++) {
        double xi = x[i], yi = y[i], ri = rmax[i] + range;
        int32_t cx = (int32_t)(xi / wx), cy = (int32_t)(yi / wy);
        if (cx > ncx - 1) cx = ncx - 1;
        if (cy > ncy - 1) cy = ncy - 1;
        if (cx < 0) cx = 0;
        if (cy < 0) cy = 0;
        int32_t rx = ncx >= 3 ? 1 : 0, ry = ncy >= 3 ? 1 : 0;
        int32_t n = 0;
        for (int32_t oy = -ry; oy <= ry; oy++)
            for (int32_t ox = -rx; ox <= rx; ox++) {
                int32_t cell = imod(cx + ox, ncx) + ncx * imod(cy + oy, ncy);
                for (int32_t q = start[cell]; q <= stop[cell]; q++) {
                    int32_t j = perm[q];
                    if (j == i) continue;
                    double dx = fcco_minimum_image(xi - x[j], Lx);
                    double dy = fcco_minimum_image(yi - y[j], Ly);
                    double lim = ri + rmax[j];
                    if (dx * dx + dy * dy <= lim * lim) {
                        n++;
                        if (n <= cap) idx[(n - 1) + (int64_t)cap * i] = j;
                    }
                }
            }
        count[i] = n;
        if (n > nmax) nmax = n;
    }
    return nmax;
}

int fcco_build_lists(int64_t n, const double *x, const double *y, const double *rmax, double Lx,
                     double Ly, double cutoff, double range, int32_t capacity, int32_t *ncx,
                     int32_t *ncy, int32_t *cellid, int32_t *perm, int32_t *start, int32_t *stop,
                     int32_t *count, int32_t *idx, int32_t *max_count) {
    int32_t nx = ncells(Lx, cutoff), ny = ncells(Ly, cutoff);
    double wx = Lx / nx, wy = Ly / ny;
    int64_t nce = (int64_t)nx * ny;
    int32_t *scratch = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nce > 0 ? nce : 1));
    if (!scratch) return fail(FCC_ERR_ARGUMENT, "out of memory");
    celllist_build(n, x, y, nx, ny, wx, wy, cellid, perm, start, stop, scratch);
    free(scratch);
    *max_count = neighbor_build(n, x, y, rmax, perm, start, stop, Lx, Ly, nx, ny, wx, wy, range,
                                capacity, idx, count);
    for (int64_t i = 0; i < n; i++) {
        cellid[i] += 1;
        perm[i] += 1;
    }
    for (int64_t c = 0; c < nce; c++) {
        start[c] += 1;
        stop[c] += 1;
    }
    for (int64_t i = 0; i < n; i++)
        for (int32_t q = 0; q < capacity; q++) {
            int64_t o = q + (int64_t)capacity * i;
            idx[o] = q < count[i] ? idx[o] + 1 : 0;
        }
    *ncx = nx;
    *ncy = ny;
    return FCC_OK;
}

/* update_neighbors!(nl, ...) with capacity regrowth, src/neighbors.jl:223-242. */
static int neighbors_rebuild(fcco_sim *s) {
    celllist_build(s->N, s->x, s->y, s->ncx, s->ncy, s->wx, s->wy, s->cellid, s->perm, s->start,
                   s->stop, s->scratch);
    for (;;) {
        int32_t nmax = neighbor_build(s->N, s->x, s->y, s->rmax, s->perm, s->start, s->stop, s->Lx,
                                      s->Ly, s->ncx, s->ncy, s->wx, s->wy, s->range + s->skin,
                                      s->cap, s->nidx, s->ncount);
        if (nmax <= s->cap) break;
        int32_t cap = 2 * nmax;
        int32_t *ni = (int32_t *)calloc((size_t)cap * (size_t)s->N, sizeof(int32_t));
        if (!ni) return fail(FCC_ERR_ARGUMENT, "out of memory");
        free(s->nidx);
        s->nidx = ni;
        s->cap = cap;
    }
    if (s->skin > 0.0) {
        memcpy(s->x0, s->x, sizeof(double) * (size_t)s->N);
        memcpy(s->y0, s->y, sizeof(double) * (size_t)s->N);
        memcpy(s->rmax0, s->rmax, sizeof(double) * (size_t)s->N);
    }
    s->valid = 1;
    s->builds++;
    return FCC_OK;
}

/* skin_flag!, src/neighbors.jl:198-205. */
static int skin_exceeded(const fcco_sim *s, int64_t i, double x, double y, double rmax) {
    double dx = fcco_minimum_image(x - s->x0[i], s->Lx);
    double dy = fcco_minimum_image(y - s->y0[i], s->Ly);
    double dr = rmax - s->rmax0[i];
    double delta = sqrt(dx * dx + dy * dy) + (dr > 0.0 ? dr : 0.0);
    return 2.0 * delta > s->skin;
}

/* update_neighbors!(sim; force), src/simulation.jl:132-142 (skin_check!
 * src/neighbors.jl:211-215). */
int fcco_update_neighbors(fcco_sim *s, int32_t force) {
    if (!force && s->valid && s->skin > 0.0) {
        int flag = 0;
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(| : flag)
        for (int64_t i = 0; i < s->N; i++) flag |= skin_exceeded(s, i, s->x[i], s->y[i], s->rmax[i]);
        if (flag) s->status[STATUS_REBUILD] = 1;
        if (s->status[STATUS_REBUILD] == 0) return FCC_OK;
    }
    int rc = neighbors_rebuild(s);
    if (rc) return rc;
    s->status[0] = s->status[1] = s->status[2] = 0;
    return FCC_OK;
}

/* ---------------------------------------------------------------- update */

/* substrate_displacement, src/dynamics.jl:84-99. */
static void substrate_displacement(const fcco_sim *s, int64_t i, const double *xi_noise,
                                   double *dx, double *dy) {
    const fcc_body_type *p = &s->types[s->type[i]];
    double fx = s->gx[i], fy = s->gy[i];
    for (int32_t q = 0; q < s->ccount[i]; q++) {
        int32_t k = s->cidx[q + (int64_t)s->ccap * i];
        if (s->types[s->type[k]].parent_drag) {
            fx += s->gx[k];
            fy += s->gy[k];
        }
    }
    double sn = sin(s->th[i]), c = cos(s->th[i]);
    double st = sqrt(2.0 * p->Dt * s->dt);
    double x1, x2;
    if (xi_noise) {
        x1 = xi_noise[3 * i];
        x2 = xi_noise[3 * i + 1];
    } else {
        philox_pair64(s->key, i + 1, (uint64_t)s->clock, 0, &x1, &x2);
    }
    *dx = s->dt * (p->v0 * c - p->mur * fx) + st * x1;
    *dy = s->dt * (p->v0 * sn - p->mur * fy) + st * x2;
}

/* displacement_kernel!, src/dynamics.jl:128-147 (noise from a 3 x N matrix
 * or generated, PhiloxNoise src/random.jl:88-100). */
static void displacement(fcco_sim *s, const double *xi_noise) {
    s->status[STATUS_UPDATE] = s->status[STATUS_REBUILD];
    if (s->status[STATUS_REBUILD] != 0) return;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t i = 0; i < s->N; i++) {
        const fcc_body_type *p = &s->types[s->type[i]];
        double d0, d1;
        if (p->parent_drag) {
            int64_t c = s->parent[i];
            double ux, uy;
            substrate_displacement(s, c, xi_noise, &ux, &uy);
            double sn = sin(s->th[i]), co = cos(s->th[i]);
            double st = sqrt(2.0 * p->Dt * s->dt);
            double x1, x2;
            if (xi_noise) {
                x1 = xi_noise[3 * i];
                x2 = xi_noise[3 * i + 1];
            } else {
                philox_pair64(s->key, i + 1, (uint64_t)s->clock, 0, &x1, &x2);
            }
            d0 = ux + s->dt * (p->v0 * co - p->mur * s->gx[i]) + st * x1;
            d1 = uy + s->dt * (p->v0 * sn - p->mur * s->gy[i]) + st * x2;
        } else {
            substrate_displacement(s, i, xi_noise, &d0, &d1);
        }
        double x3;
        if (xi_noise) {
            x3 = xi_noise[3 * i + 2];
        } else {
            double x4;
            philox_pair64(s->key, i + 1, (uint64_t)s->clock, 1, &x3, &x4);
        }
        s->disp[3 * i] = d0;
        s->disp[3 * i + 1] = d1;
        s->disp[3 * i + 2] = sqrt(2.0 * p->Dr * s->dt) * x3;
    }
}

/* euler_maruyama_kernel!, src/dynamics.jl:180-203. `geom`: update lab/rmax;
 * `skin`: fused skin check; `clock`: tick. */
static void euler_maruyama(fcco_sim *s, int frame, int geom, int skin, int clock) {
    if (s->status[STATUS_UPDATE] != 0) return;
    int K = s->K, nc = s->nc;
    int flag = 0;
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(| : flag)
    for (int64_t i = 0; i < s->N; i++) {
        const fcc_body_type *p = &s->types[s->type[i]];
        if (frame) body_frame(s, i, s->th[i]);
        double thn = s->th[i] - s->dt * p->mutheta * s->gth[i] + s->disp[3 * i + 2];
        s->th[i] = thn;
        double *c = s->coef + nc * i;
        const double *gc = s->gcoef + nc * i;
        for (int r = 1; r < nc; r++) c[r] -= s->dt * p->muab * gc[r];
        enforce_shape(c, K);
        int32_t sx, sy;
        double xw = fcco_wrap(s->x[i] + s->disp[3 * i], s->Lx, &sx);
        double yw = fcco_wrap(s->y[i] + s->disp[3 * i + 1], s->Ly, &sy);
        s->x[i] = xw;
        s->y[i] = yw;
        s->imgx[i] += sx;
        s->imgy[i] += sy;
        if (geom) lab_frame(s, i, thn);
        if (skin && skin_exceeded(s, i, xw, yw, s->rmax[i])) flag = 1;
    }
    if (flag) s->status[STATUS_REBUILD] = 1;
    if (clock) s->clock += 1;
}

/* ------------------------------------------------------------------ bonds */

/* bond_gradient(::HarmonicBond, Δ, ti, tj), src/energies/bonds.jl:25-29:
 * ∂E/∂Δ = k (1 - r0/r) Δ; at r = 0, k Δ (r0 = 0) or zero (r0 > 0). */
static void harmonic_bond_gradient(const fcc_term *t, double dx, double dy, double g[2]) {
    const double r = sqrt(dx * dx + dy * dy);
    const double k = t->p[0], r0 = t->p[1];
    const double f = r > 0.0 ? k * (1.0 - r0 / r) : (r0 == 0.0 ? k : 0.0);
    g[0] = f * dx;
    g[1] = f * dy;
}

/* bond_kernel!, src/energies/bonds.jl:98-110, with the right fold over the
 * terms of _bond_gradient (:80-85): gx, gy += sum over partners j of
 * ∂E/∂Δ at Δ = minimum image of r_i - r_j. */
static void bond_gradients(fcco_sim *s) {
    if (s->nbondt == 0 || !s->bcount) return;
    const int64_t N = s->N;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t i = 0; i < N; i++) {
        double gi0 = 0.0, gi1 = 0.0;
        for (int32_t q = 0; q < s->bcount[i]; q++) {
            const int32_t j = s->bidx[q + (int64_t)s->bcap * i];
            const double dx = fcco_minimum_image(s->x[i] - s->x[j], s->Lx);
            const double dy = fcco_minimum_image(s->y[i] - s->y[j], s->Ly);
            double a0 = 0.0, a1 = 0.0;
            for (int t = s->nbondt - 1; t >= 0; t--) {
                double g[2] = {0.0, 0.0};
                if (applies(&s->bondt[t], s->type[i]) && applies(&s->bondt[t], s->type[j]))
                    harmonic_bond_gradient(&s->bondt[t], dx, dy, g);
                a0 = g[0] + a0;
                a1 = g[1] + a1;
            }
            gi0 += a0;
            gi1 += a1;
        }
        s->gx[i] += gi0;
        s->gy[i] += gi1;
    }
}

static int cmp_i64(const void *a, const void *b) {
    const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* BondList(bonds, N), src/energies/bonds.jl:54-72: checks, then every bond
 * (i, j) appended to the columns of i and of j in the order given. */
int fcco_set_bonds(fcco_sim *s, int64_t n, const int32_t *pairs) {
    const int64_t N = s->N;
    if (n < 0 || (n > 0 && !pairs)) return fail(FCC_ERR_ARGUMENT, "bad bond arguments");
    int32_t *count = calloc((size_t)N, sizeof(int32_t));
    if (!count) return fail(FCC_ERR_ARGUMENT, "out of memory");
    for (int64_t q = 0; q < n; q++) {
        const int32_t i = pairs[2 * q], j = pairs[2 * q + 1];
        if (i < 1 || i > N || j < 1 || j > N) {
            free(count);
            return fail(FCC_ERR_ARGUMENT, "bond (%d, %d) refers to a body outside 1:%lld", i, j, (long long)N);
        }
        if (i == j) {
            free(count);
            return fail(FCC_ERR_ARGUMENT, "bond (%d, %d) connects a body to itself", i, j);
        }
        count[i - 1]++;
        count[j - 1]++;
    }
    /* duplicates: equal (min, max) keys, found by sorting */
    int64_t *key = malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    if (!key) {
        free(count);
        return fail(FCC_ERR_ARGUMENT, "out of memory");
    }
    for (int64_t q = 0; q < n; q++) {
        const int64_t a = pairs[2 * q], b = pairs[2 * q + 1];
        key[q] = a < b ? a * (N + 1) + b : b * (N + 1) + a;
    }
    qsort(key, (size_t)n, sizeof(int64_t), cmp_i64);
    for (int64_t q = 1; q < n; q++)
        if (key[q] == key[q - 1]) {
            const long long a = key[q] / (N + 1), b = key[q] % (N + 1);
            free(key);
            free(count);
            return fail(FCC_ERR_ARGUMENT, "duplicate bond (%lld, %lld)", a, b);
        }
    free(key);
    int32_t cap = 1;
    for (int64_t i = 0; i < N; i++)
        if (count[i] > cap) cap = count[i];
    int32_t *idx = calloc((size_t)cap * (size_t)N, sizeof(int32_t));
    if (!idx) {
        free(count);
        return fail(FCC_ERR_ARGUMENT, "out of memory");
    }
    memset(count, 0, sizeof(int32_t) * (size_t)N);
    for (int64_t q = 0; q < n; q++) {
        const int32_t ab[2][2] = {{pairs[2 * q] - 1, pairs[2 * q + 1] - 1}, {pairs[2 * q + 1] - 1, pairs[2 * q] - 1}};
        for (int e = 0; e < 2; e++) {
            const int32_t a = ab[e][0], b = ab[e][1];
            idx[count[a] + (int64_t)cap * a] = b;
            count[a]++;
        }
    }
    free(s->bidx);
    free(s->bcount);
    s->bidx = idx;
    s->bcount = count;
    s->bcap = cap;
    return FCC_OK;
}

/* --------------------------------------------------------------- stepping */

/* compute_gradients!, src/simulation.jl:150-158. */
int fcco_compute_gradients(fcco_sim *s, int32_t frame) {
    if (s->nbody > 0) {
        body_terms(s);
    } else {
        int nth = s->nth;
        for (int64_t i = 0; i < s->N; i++) {
            for (int k = 0; k < nth; k++) {
                s->GR[k + nth * i] = 0.0;
                s->GD[k + nth * i] = 0.0;
                s->GDD[k + nth * i] = 0.0;
            }
            s->gx[i] = s->gy[i] = s->gth[i] = s->Ebody[i] = 0.0;
        }
    }
    if (s->ncontact > 0) {
        if (!s->valid) return fail(FCC_ERR_STATE, "neighbour list not built");
        contact_gradients(s);
    }
    bond_gradients(s);
    pullback(s, frame);
    return FCC_OK;
}

/* advance!, src/simulation.jl:181-191, with integrate! src/dynamics.jl:220-232. */
static void advance(fcco_sim *s) {
    fcco_compute_gradients(s, 0);
    int skin = s->ncontact > 0 && s->valid && s->skin > 0.0;
    displacement(s, NULL);
    euler_maruyama(s, 1, 1, skin, 1);
    synthesize(s);
}

/* step!, src/simulation.jl:165-173. */
int fcco_step(fcco_sim *s, int64_t n) {
    for (int64_t q = 0; q < n; q++) {
        if (s->ncontact > 0) {
            int rc = fcco_update_neighbors(s, 0);
            if (rc) return rc;
        }
        advance(s);
        s->t += s->dt;
        s->step += 1;
    }
    return FCC_OK;
}

int fcco_set_center_gradients(fcco_sim *s, const double *gx, const double *gy) {
    memcpy(s->gx, gx, sizeof(double) * (size_t)s->N);
    memcpy(s->gy, gy, sizeof(double) * (size_t)s->N);
    return FCC_OK;
}

/* integrate!(bodies, grad, disp, ξ, children, types, dt, box, basis) with the
 * defaults geom = skin = clock = status = nothing, frame = false. */
int fcco_integrate_noise(fcco_sim *s, const double *xi) {
    int32_t st[3] = {s->status[0], s->status[1], s->status[2]};
    s->status[0] = s->status[1] = s->status[2] = 0;
    displacement(s, xi);
    euler_maruyama(s, 0, 0, 0, 0);
    s->status[0] = st[0];
    s->status[1] = st[1];
    s->status[2] = st[2];
    return FCC_OK;
}

/* ------------------------------------------------------------- lifecycle */

static void *xcalloc(size_t n, size_t sz, int *bad) {
    void *p = calloc(n ? n : 1, sz);
    if (!p) *bad = 1;
    return p;
}

int fcco_destroy(fcco_sim *s) {
    if (!s) return FCC_OK;
    void *ptrs[] = {s->C, s->D, s->DD, s->cs, s->x, s->y, s->th, s->coef, s->R0, s->imgx, s->imgy,
                    s->type, s->parent, s->lab, s->labR0, s->R, s->dR, s->ddR, s->rmax, s->GR,
                    s->GD, s->GDD, s->P, s->gx, s->gy, s->gth, s->glab, s->gcoef, s->Ebody,
                    s->cellid, s->perm, s->start, s->stop, s->scratch, s->nidx, s->ncount, s->x0,
                    s->y0, s->rmax0, s->cidx, s->ccount, s->disp, s->bidx, s->bcount};
    for (size_t q = 0; q < sizeof ptrs / sizeof ptrs[0]; q++) free(ptrs[q]);
    free(s);
    return FCC_OK;
}

static int load_bodies(fcco_sim *s, const fcc_bodies *b) {
    int64_t N = s->N;
    int nc = s->nc;
    for (int64_t i = 0; i < N; i++) {
        s->x[i] = b->x[i];
        s->y[i] = b->y[i];
        s->th[i] = b->theta[i];
        s->imgx[i] = b->imgx ? b->imgx[i] : 0;
        s->imgy[i] = b->imgy ? b->imgy[i] : 0;
        s->R0[i] = b->R0[i];
        s->type[i] = (b->type ? b->type[i] : 1) - 1;
        s->parent[i] = (b->parent ? b->parent[i] : (int32_t)(i + 1)) - 1;
        for (int c = 0; c < nc; c++) s->coef[c + (int64_t)nc * i] = b->coef[c + (int64_t)nc * i];
    }
    return FCC_OK;
}

/* Simulation(...) validation, src/simulation.jl:71-118, basis.jl:36-38,
 * neighbors.jl:123. */
static int validate(const fcc_config *cfg, const fcc_bodies *b) {
    if (cfg->n_bodies < 1) return fail(FCC_ERR_ARGUMENT, "need at least one body");
    if (cfg->n_modes < 1) return fail(FCC_ERR_ARGUMENT, "maxmode must be ≥ 2");
    if (cfg->n_theta <= 2 * (cfg->n_modes + 1)) return fail(FCC_ERR_ARGUMENT, "nθ must exceed 2·maxmode");
    if (cfg->n_types < 1 || cfg->n_types > FCC_MAX_TYPES) return fail(FCC_ERR_UNSUPPORTED, "1..%d body types", FCC_MAX_TYPES);
    if (cfg->n_terms < 0 || cfg->n_terms > FCC_MAX_TERMS) return fail(FCC_ERR_UNSUPPORTED, "at most %d energy terms", FCC_MAX_TERMS);
    if (cfg->skin < 0.0) return fail(FCC_ERR_ARGUMENT, "skin must be non-negative");
    if (!b || !b->x || !b->y || !b->theta || !b->coef || !b->R0) return fail(FCC_ERR_DIMENSION, "missing body arrays");
    int64_t N = cfg->n_bodies;
    for (int64_t i = 0; i < N; i++) {
        int32_t t = b->type ? b->type[i] : 1;
        if (t < 1 || t > cfg->n_types) return fail(FCC_ERR_ARGUMENT, "body type without BodyType entry");
    }
    for (int64_t i = 0; i < N; i++) {
        int32_t p = b->parent ? b->parent[i] : (int32_t)(i + 1);
        if (p < 1 || p > N) return fail(FCC_ERR_ARGUMENT, "body %lld has parent %d outside 1:%lld", (long long)(i + 1), p, (long long)N);
        int32_t t = b->type ? b->type[i] : 1;
        if (!cfg->types[t - 1].parent_drag) continue;
        if (p == i + 1) return fail(FCC_ERR_ARGUMENT, "body %lld has drag = :parent but is its own parent", (long long)(i + 1));
        int32_t tp = b->type ? b->type[p - 1] : 1;
        if (cfg->types[tp - 1].parent_drag) return fail(FCC_ERR_ARGUMENT, "body %lld has drag = :parent, but so has its parent %d (one level only)", (long long)(i + 1), p);
    }
    for (int q = 0; q < cfg->n_terms; q++) {
        int k = cfg->terms[q].kind;
        if (k != FCC_TERM_TENSION && k != FCC_TERM_PERIMETER_ELASTICITY && k != FCC_TERM_BENDING &&
            k != FCC_TERM_OVERLAP && k != FCC_TERM_LINEAR_REPULSION && k != FCC_TERM_HARMONIC_BOND)
            return fail(FCC_ERR_ARGUMENT, "energy terms must be BodyTerms, ContactTerms or BondTerms");
    }
    return FCC_OK;
}

int fcco_create(const fcc_config *cfg, const fcc_bodies *b, fcco_sim **out) {
    int rc = validate(cfg, b);
    if (rc) return rc;
    fcco_sim *s = (fcco_sim *)calloc(1, sizeof(fcco_sim));
    if (!s) return fail(FCC_ERR_ARGUMENT, "out of memory");
    int64_t N = cfg->n_bodies;
    s->N = N;
    s->K = cfg->n_modes;
    s->nc = 2 * s->K + 1;
    s->nth = cfg->n_theta;
    s->ntypes = cfg->n_types;
    s->Lx = cfg->Lx;
    s->Ly = cfg->Ly;
    s->dt = cfg->dt;
    s->skin = cfg->skin;
    memcpy(s->types, cfg->types, sizeof s->types);
    philox_key(cfg->seed, s->key);
    double range = 0.0;
    for (int q = 0; q < cfg->n_terms; q++) {
        const fcc_term *t = &cfg->terms[q];
        if (t->kind == FCC_TERM_OVERLAP || t->kind == FCC_TERM_LINEAR_REPULSION)
            s->contact[s->ncontact++] = *t;
        else if (t->kind == FCC_TERM_HARMONIC_BOND)
            s->bondt[s->nbondt++] = *t;
        else
            s->body[s->nbody++] = *t;
    }
    s->range = range; /* interaction_range of Overlap and LinearRepulsion is 0 */
    int bad = 0;
    size_t n = (size_t)N, nth = (size_t)s->nth, nc = (size_t)s->nc;
    s->C = xcalloc(nth * nc, sizeof(double), &bad);
    s->D = xcalloc(nth * nc, sizeof(double), &bad);
    s->DD = xcalloc(nth * nc, sizeof(double), &bad);
    s->cs = xcalloc(2 * nth, sizeof(double), &bad);
    s->x = xcalloc(n, sizeof(double), &bad);
    s->y = xcalloc(n, sizeof(double), &bad);
    s->th = xcalloc(n, sizeof(double), &bad);
    s->coef = xcalloc(n * nc, sizeof(double), &bad);
    s->R0 = xcalloc(n, sizeof(double), &bad);
    s->imgx = xcalloc(n, sizeof(int32_t), &bad);
    s->imgy = xcalloc(n, sizeof(int32_t), &bad);
    s->type = xcalloc(n, sizeof(int32_t), &bad);
    s->parent = xcalloc(n, sizeof(int32_t), &bad);
    s->lab = xcalloc(n * nc, sizeof(double), &bad);
    s->labR0 = xcalloc(n * nc, sizeof(double), &bad);
    s->R = xcalloc(n * nth, sizeof(double), &bad);
    s->dR = xcalloc(n * nth, sizeof(double), &bad);
    s->ddR = xcalloc(n * nth, sizeof(double), &bad);
    s->rmax = xcalloc(n, sizeof(double), &bad);
    s->GR = xcalloc(n * nth, sizeof(double), &bad);
    s->GD = xcalloc(n * nth, sizeof(double), &bad);
    s->GDD = xcalloc(n * nth, sizeof(double), &bad);
    s->P = xcalloc(n * nth, sizeof(double), &bad);
    s->gx = xcalloc(n, sizeof(double), &bad);
    s->gy = xcalloc(n, sizeof(double), &bad);
    s->gth = xcalloc(n, sizeof(double), &bad);
    s->glab = xcalloc(n * nc, sizeof(double), &bad);
    s->gcoef = xcalloc(n * nc, sizeof(double), &bad);
    s->Ebody = xcalloc(n, sizeof(double), &bad);
    s->disp = xcalloc(3 * n, sizeof(double), &bad);
    if (bad) {
        fcco_destroy(s);
        return fail(FCC_ERR_ARGUMENT, "out of memory");
    }
    basis_init(s);
    load_bodies(s, b);
    /* cutoff, src/simulation.jl:90; CellList src/neighbors.jl:27-33 */
    double maxR0 = 0.0;
    for (int64_t i = 0; i < N; i++)
        if (s->R0[i] > maxR0) maxR0 = s->R0[i];
    double cutoff = 2.0 * maxR0 * (1.0 + sqrt(2.0 / 3.0)) + s->range + s->skin;
    s->ncx = ncells(s->Lx, cutoff);
    s->ncy = ncells(s->Ly, cutoff);
    s->wx = s->Lx / s->ncx;
    s->wy = s->Ly / s->ncy;
    size_t nce = (size_t)s->ncx * (size_t)s->ncy;
    s->cellid = xcalloc(n, sizeof(int32_t), &bad);
    s->perm = xcalloc(n, sizeof(int32_t), &bad);
    s->start = xcalloc(nce, sizeof(int32_t), &bad);
    s->stop = xcalloc(nce, sizeof(int32_t), &bad);
    s->scratch = xcalloc(nce, sizeof(int32_t), &bad);
    /* NeighborList, src/neighbors.jl:121-127 (capacity 32) */
    s->cap = 32;
    s->nidx = xcalloc((size_t)s->cap * n, sizeof(int32_t), &bad);
    s->ncount = xcalloc(n, sizeof(int32_t), &bad);
    s->x0 = xcalloc(n, sizeof(double), &bad);
    s->y0 = xcalloc(n, sizeof(double), &bad);
    s->rmax0 = xcalloc(n, sizeof(double), &bad);
    /* ChildList, src/bodies.jl:121-135 */
    s->ccount = xcalloc(n, sizeof(int32_t), &bad);
    if (bad) {
        fcco_destroy(s);
        return fail(FCC_ERR_ARGUMENT, "out of memory");
    }
    int32_t cmax = 0;
    for (int64_t k = 0; k < N; k++)
        if (s->parent[k] != k) s->ccount[s->parent[k]]++;
    for (int64_t k = 0; k < N; k++)
        if (s->ccount[k] > cmax) cmax = s->ccount[k];
    s->ccap = cmax > 1 ? cmax : 1;
    s->cidx = xcalloc((size_t)s->ccap * n, sizeof(int32_t), &bad);
    if (bad) {
        fcco_destroy(s);
        return fail(FCC_ERR_ARGUMENT, "out of memory");
    }
    memset(s->ccount, 0, sizeof(int32_t) * n);
    for (int64_t k = 0; k < N; k++) {
        int32_t p = s->parent[k];
        if (p == k) continue;
        s->cidx[s->ccount[p] + (int64_t)s->ccap * p] = (int32_t)k;
        s->ccount[p]++;
    }
    /* enforce_shape!(dev, basis); update_geometry!, src/simulation.jl:102-103 */
    for (int64_t i = 0; i < N; i++) enforce_shape(s->coef + nc * i, s->K);
    update_geometry(s);
    *out = s;
    return FCC_OK;
}

int fcco_set_bodies(fcco_sim *s, const fcc_bodies *b) {
    load_bodies(s, b);
    for (int64_t i = 0; i < s->N; i++) enforce_shape(s->coef + (int64_t)s->nc * i, s->K);
    update_geometry(s);
    s->valid = 0;
    return FCC_OK;
}

int fcco_get_bodies(fcco_sim *s, fcc_bodies *b) {
    int64_t N = s->N;
    int nc = s->nc;
    for (int64_t i = 0; i < N; i++) {
        if (b->x) b->x[i] = s->x[i];
        if (b->y) b->y[i] = s->y[i];
        if (b->theta) b->theta[i] = s->th[i];
        if (b->imgx) b->imgx[i] = s->imgx[i];
        if (b->imgy) b->imgy[i] = s->imgy[i];
        if (b->R0) b->R0[i] = s->R0[i];
        if (b->type) b->type[i] = s->type[i] + 1;
        if (b->parent) b->parent[i] = s->parent[i] + 1;
        if (b->coef)
            for (int c = 0; c < nc; c++) b->coef[c + (int64_t)nc * i] = s->coef[c + (int64_t)nc * i];
    }
    return FCC_OK;
}

int fcco_set_dt(fcco_sim *s, double dt) {
    s->dt = dt;
    return FCC_OK;
}

int fcco_get_gradients(fcco_sim *s, double *gx, double *gy, double *gth, double *glab,
                       double *gcoef, double *Ebody) {
    size_t n = (size_t)s->N, nc = (size_t)s->nc;
    if (gx) memcpy(gx, s->gx, n * sizeof(double));
    if (gy) memcpy(gy, s->gy, n * sizeof(double));
    if (gth) memcpy(gth, s->gth, n * sizeof(double));
    if (glab) memcpy(glab, s->glab, n * nc * sizeof(double));
    if (gcoef) memcpy(gcoef, s->gcoef, n * nc * sizeof(double));
    if (Ebody) memcpy(Ebody, s->Ebody, n * sizeof(double));
    return FCC_OK;
}

int fcco_get_contour_gradients(fcco_sim *s, double *P, double *GR, double *GD, double *GDD) {
    size_t m = (size_t)s->N * (size_t)s->nth * sizeof(double);
    if (P) memcpy(P, s->P, m);
    if (GR) memcpy(GR, s->GR, m);
    if (GD) memcpy(GD, s->GD, m);
    if (GDD) memcpy(GDD, s->GDD, m);
    return FCC_OK;
}

int fcco_get_geometry(fcco_sim *s, double *lab, double *rmax, double *R, double *dR, double *ddR) {
    size_t n = (size_t)s->N, m = n * (size_t)s->nth * sizeof(double);
    if (lab) memcpy(lab, s->lab, n * (size_t)s->nc * sizeof(double));
    if (rmax) memcpy(rmax, s->rmax, n * sizeof(double));
    if (R) memcpy(R, s->R, m);
    if (dR) memcpy(dR, s->dR, m);
    if (ddR) memcpy(ddR, s->ddR, m);
    return FCC_OK;
}

/* perimeters / areas / shape_index, src/observables.jl:6-18. */
int fcco_shape_index(fcco_sim *s, double *per, double *area, double *p) {
    int nth = s->nth;
    for (int64_t i = 0; i < s->N; i++) {
        double a = 0.0, l = 0.0;
        for (int k = 0; k < nth; k++) {
            double R = s->R[k + nth * i], dR = s->dR[k + nth * i];
            l += sqrt(R * R + dR * dR);
            a += R * R;
        }
        l *= s->dphi;
        a *= s->dphi / 2.0;
        if (per) per[i] = l;
        if (area) area[i] = a;
        if (p) p[i] = l / sqrt(a);
    }
    return FCC_OK;
}

/* contours(sim), src/observables.jl:36-44: X = x' .+ R .* cos.(φ),
 * Y = y' .+ R .* sin.(φ) with φ = node_angle (src/basis.jl:73). */
int fcco_contours(fcco_sim *s, double *X, double *Y) {
    const int nth = s->nth;
    for (int64_t i = 0; i < s->N; i++)
        for (int k = 0; k < nth; k++) {
            const double R = s->R[k + nth * i];
            X[k + nth * i] = s->x[i] + R * s->cs[2 * k];
            Y[k + nth * i] = s->y[i] + R * s->cs[2 * k + 1];
        }
    return FCC_OK;
}

int fcco_get_celllist(fcco_sim *s, int32_t *ncx, int32_t *ncy, int32_t *cellid, int32_t *perm,
                      int32_t *start, int32_t *stop) {
    if (ncx) *ncx = s->ncx;
    if (ncy) *ncy = s->ncy;
    int64_t nce = (int64_t)s->ncx * s->ncy;
    for (int64_t i = 0; i < s->N; i++) {
        if (cellid) cellid[i] = s->cellid[i] + 1;
        if (perm) perm[i] = s->perm[i] + 1;
    }
    for (int64_t c = 0; c < nce; c++) {
        if (start) start[c] = s->start[c] + 1;
        if (stop) stop[c] = s->stop[c] + 1;
    }
    return FCC_OK;
}

int fcco_get_neighbors(fcco_sim *s, int32_t *capacity, int32_t *count, int32_t *idx) {
    if (capacity) *capacity = s->cap;
    if (count) memcpy(count, s->ncount, (size_t)s->N * sizeof(int32_t));
    if (idx)
        for (int64_t i = 0; i < s->N; i++)
            for (int32_t q = 0; q < s->cap; q++) {
                int64_t o = q + (int64_t)s->cap * i;
                idx[o] = q < s->ncount[i] ? s->nidx[o] + 1 : 0;
            }
    return FCC_OK;
}

int fcco_set_neighbor_capacity(fcco_sim *s, int32_t capacity) {
    if (capacity < 1) return fail(FCC_ERR_ARGUMENT, "capacity must be positive");
    int32_t *ni = (int32_t *)calloc((size_t)capacity * (size_t)s->N, sizeof(int32_t));
    if (!ni) return fail(FCC_ERR_ARGUMENT, "out of memory");
    free(s->nidx);
    s->nidx = ni;
    s->cap = capacity;
    s->valid = 0;
    return FCC_OK;
}

int fcco_get_counters(fcco_sim *s, int64_t *step, double *t, int64_t *clock, int64_t *builds) {
    if (step) *step = s->step;
    if (t) *t = s->t;
    if (clock) *clock = s->clock;
    if (builds) *builds = s->builds;
    return FCC_OK;
}
