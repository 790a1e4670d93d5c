// fcc_device.hpp — device-side building blocks of the gfx950 time step:
// kernel parameter blocks, Philox4x32-10, periodic-box helpers, the built-in
// body terms and the off-grid contour radius.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fcc.h"

namespace fcc {

constexpr int MAXB = 4;  // native body terms per simulation
// Nodes per contour of the one-lane-per-body forces kernels (64-bit node
// masks, per-lane LDS rows); larger contours take k_forces_wide.
constexpr int FUSED_MAX_THETA = 64;

// Device run state (src/utils.jl:50-78): `rebuild` is STATUS_REBUILD, set by
// the fused skin check; `update` is STATUS_UPDATE, the copy of the freeze
// condition taken at the start of a step, which gates the update kernel;
// `overflow` holds the largest neighbour count of the last build; `clock`
// counts the executed steps and is the Philox draw index. `stop` is raised by
// the update kernel once `clock` reaches `target` (ensemble runs give every
// replica its own step budget); a raised `rebuild` or `stop` freezes the
// following steps. `builds` counts the list builds run on the device side of
// an ensemble (the host compares it with the builds it has issued).
struct DevStatus {
    int32_t rebuild, update, overflow, stop;
    int64_t clock;
    int64_t target;
    int32_t builds;
    int32_t cur;  // ensembles of one-kernel steps: parity of the current state's buffers
    // One-kernel steps, by the parity q of the buffers a step reads: the
    // step's clock (cslot[q], passed on by the step before), and the clock
    // of the last step before it whose bodies tripped the skin check
    // (tstep[q] = clock - 1 freezes the step).
    int64_t cslot[2], tstep[2];
};

// Status prologue of the first kernel of a step: block 0 mirrors the freeze
// condition into `update` and ticks the clock of an executed step. Returns
// whether this step is frozen (uniform over the launch).
__device__ __forceinline__ bool step_frozen(DevStatus *st, int64_t blk, int lane) {
    const int32_t fr = st->rebuild | st->stop;
    if (blk == 0 && lane == 0) {
        st->update = fr;
        if (fr == 0) st->clock += 1;
    }
    return fr != 0;
}

// Tile of workgroup b in a launch of n workgroups, grouped by XCD. The
// dispatcher places workgroup b on XCD b mod 8; XCD x takes the contiguous
// tiles [x·q + min(x, r), … + q + (x < r)) with q = n / 8, r = n mod 8 (a
// bijection on [0, n)), so the neighbours a tile gathers (the lattice rows on
// either side) were mostly fetched into the same XCD's L2 by nearby tiles.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t n) {
    const int64_t x = b & 7, t = b >> 3, q = n >> 3, r = n & 7;
    return x * q + (x < r ? x : r) + t;
}

// Status check of the forces kernel: use_status 1 = the kernel opens the step
// (step_frozen); 2 = a gate kernel already took the step's freeze decision
// into `update` (ensemble steps, whose lists are rebuilt concurrently).
// One-kernel steps open with step_open instead.
__device__ __forceinline__ bool forces_frozen(int use_status, DevStatus *st, int64_t blk, int lane) {
    if (use_status == 1) return step_frozen(st, blk, lane);
    if (use_status == 2) return st->update != 0;
    return false;
}

// Status prologue of a one-kernel step reading the buffers of parity q. Every
// workgroup reads only fields that no workgroup of the launch writes, so the
// decision is uniform: the step is frozen by a raised rebuild flag, a reached
// step budget or a skin trip of the step before (tstep[q] = clock - 1; that
// step wrote the marker while this launch's trips go to tstep[1 - q]).
// Workgroup 0 passes the clock on to the next step (cslot[1 - q]) and to
// `clock`, or, in a step frozen by a trip, raises the rebuild flag so that
// the freeze holds for the steps after it. `timing` (state-restoring timing
// series) never freezes. Returns whether the step is frozen; clk is its clock
// (the Philox draw index).
__device__ __forceinline__ bool step_open(DevStatus *st, int q, bool timing, int64_t blk, int lane,
                                          int64_t &clk) {
    const int64_t c = st->cslot[q];
    const bool trip = st->tstep[q] == c - 1;
    const bool fr = !timing && (st->rebuild != 0 || st->stop != 0 || c >= st->target || trip);
    if (blk == 0 && lane == 0) {
        if (!fr) {
            st->cslot[q ^ 1] = c + 1;
            st->clock = c + 1;
            st->update = 0;
        } else if (trip) {
            st->rebuild = 1;
        }
    }
    clk = c;
    return fr;
}

// Scalar parameters of one simulation, passed by value to every kernel.
template <typename T>
struct Params {
    int64_t N;
    int K, nc, ldc, nth;
    T dphi;
    double Lx, Ly;
    // body terms (Tension, PerimeterElasticity, Bending) with OnTypes masks
    int nbody;
    int bkind[MAXB];
    uint32_t bmask[MAXB];
    T bp0[MAXB], bp1[MAXB];
    int need_dd;  // a term uses R'' (Bending)
    // contact: Σ over Overlap terms of applies(ti)·J, by (related, ti, tj)
    int ncontact;
    T Jc[2][FCC_MAX_TYPES][FCC_MAX_TYPES];
    T range;
    // body types
    T v0[FCC_MAX_TYPES], Dr[FCC_MAX_TYPES], Dt[FCC_MAX_TYPES], mur[FCC_MAX_TYPES],
        muth[FCC_MAX_TYPES], muab[FCC_MAX_TYPES];
    int one_type;  // a single body type: the update takes its parameters without a select
    // noise amplitudes √(2·Dt·dt), √(2·Dr·dt) per type, formed on the host in
    // T as the kernels would ((2·D)·dt, then the correctly rounded root)
    T sDt[FCC_MAX_TYPES], sDr[FCC_MAX_TYPES];
    int pdrag[FCC_MAX_TYPES];
    T dt;
    uint32_t key0, key1;
    T skin;
    int skin_check;  // fused Verlet skin check in the update kernel
    // Overlap couplings as classes of equal nonzero J: ccls[related][ti][tj]
    // is the class (-1: zero coupling), Jcls its value.
    int ncls;
    int ccls[2][FCC_MAX_TYPES][FCC_MAX_TYPES];
    T Jcls[2 * FCC_MAX_TYPES * FCC_MAX_TYPES];
    int small_box;  // a node offset may need a second minimum image
    double invLx, invLy;  // 1/L: minimum image of neighbour offsets outside small boxes
    int cls_uniform;      // every pair of types (related or not) couples in class 0
    T krep[FCC_MAX_TYPES];  // LinearRepulsion: Σ k of the terms applying to body type ti
    // HarmonicBond terms (k, r0) with OnTypes masks, in the order given
    int nbondt;
    uint32_t bdmask[FCC_MAX_TERMS];
    T bdk[FCC_MAX_TERMS], bdr0[FCC_MAX_TERMS];
};

// Device buffers of one simulation. Per-body coefficient rows are stored
// body-major with stride ldc (AoS) so a neighbour's row is one contiguous
// read; the neighbour list is slot-major (nidx[q*N + i]).
template <typename T>
struct Ptrs {
    double *x, *y;
    int32_t *imgx, *imgy;
    T *th, *coef, *R0;
    int32_t *type, *parent;
    T *lab, *rmax;
    T *rmin;  // lower bound R0 (a0 - Σ √(an² + bn²)) of the contour radius
    T *glab, *gx, *gy, *gth, *Ebody, *gcoef;
    T *P, *GR, *GD, *GDD;  // debug (may be null)
    int32_t *nidx, *ncount;
    int32_t cap;
    double *x0, *y0;
    T *rmax0;
    // one-kernel steps: the buffers the step writes the new state into
    // (positions, lab rows, radius bounds); the kernels read x, y, lab, rmax,
    // rmin of the current state, which neighbours in other workgroups still
    // read while a body is updated
    double *xn, *yn;
    T *labn, *rmaxn, *rminn;
    int32_t *cidx, *ccount;
    int32_t ccap;
    // parent drag: θ at the start of the step, stored by the forces kernels
    // (which do not write θ), so that a child's update reads its parent's
    // orientation while the parent's own update overwrites θ (null without
    // parent drag)
    T *th0;
    int32_t *bidx, *bcount;  // bond partners, slot-major (bidx[q*N + i]); null without bonds
    const T *S;   // [3][nth][ldc]: C, D, DD rows
    const T *cs;  // [nth][2]: cos φk, sin φk
    const T *Sc;  // [3][nth][ldc]: C, D, DD rows in coefficient order [a0, cos.., sin..]
    DevStatus *st;
};

#define FCC_CONST __attribute__((address_space(4)))

__device__ __forceinline__ void fsincos(double x, double *s, double *c) { sincos(x, s, c); }
__device__ __forceinline__ void fsincos(float x, float *s, float *c) { sincosf(x, s, c); }

template <typename T>
__device__ __forceinline__ const FCC_CONST T *as_const(const T *p) {
    return (const FCC_CONST T *)p;
}

// One workgroup of an ensemble launch: replica `rep` (-1 = padding of the
// grid) and the workgroup's index within that replica's own grid.
struct Tile {
    int32_t rep, blk;
};

// Element r of a table the host writes between launches only, read through
// the scalar cache (uniform address) word by word.
template <typename P>
__device__ __forceinline__ P load_uniform(const P *tab, int64_t r) {
    static_assert(sizeof(P) % 4 == 0, "table entries are whole words");
    P out;
    const FCC_CONST uint32_t *src = (const FCC_CONST uint32_t *)(tab + r);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
    for (size_t q = 0; q < sizeof(P) / 4; q++) dst[q] = src[q];
    return out;
}

// ------------------------------------------------------------------ Philox

// Philox4x32-10 (Salmon et al., SC'11), src/random.jl:10-36.
__device__ __forceinline__ void philox4x32(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3,
                                           uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// uniform_open + box_muller (src/random.jl:42-51) for T = float / double.
__device__ __forceinline__ void box_muller(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3,
                                           double &z1, double &z2) {
    double u1 = ((double)((((uint64_t)a0 << 32) | a1) >> 12) + 0.5) * 0x1p-52;
    double u2 = ((double)((((uint64_t)a2 << 32) | a3) >> 12) + 0.5) * 0x1p-52;
    double r = sqrt(-2.0 * log(u1));
    double s, c;
    sincos(2.0 * 3.141592653589793 * u2, &s, &c);
    z1 = r * c;
    z2 = r * s;
}

__device__ __forceinline__ void box_muller(uint32_t a0, uint32_t, uint32_t a2, uint32_t,
                                           float &z1, float &z2) {
    float u1 = ((float)(a0 >> 9) + 0.5f) * 0x1p-23f;
    float u2 = ((float)(a2 >> 9) + 0.5f) * 0x1p-23f;
    float r = sqrtf(-2.0f * logf(u1));
    float s, c;
    sincosf(2.0f * 3.14159274f * u2, &s, &c);
    z1 = r * c;
    z2 = r * s;
}

// philox_pair(T, key, i, n, block), src/random.jl:75-78 (i 1-based).
template <typename T>
__device__ __forceinline__ void philox_pair(uint32_t k0, uint32_t k1, uint32_t i1, uint64_t n,
                                            uint32_t block, T &z1, T &z2) {
    uint32_t c0 = i1, c1 = (uint32_t)n, c2 = (uint32_t)(n >> 32), c3 = block;
    philox4x32(c0, c1, c2, c3, k0, k1);
    box_muller(c0, c1, c2, c3, z1, z2);
}

// --------------------------------------------------------------------- box

// minimum_image (src/box.jl:14), round ties-to-even.
template <typename T>
__device__ __forceinline__ T min_image(T d, T L) {
#pragma clang fp contract(off)
    return d - L * rint(d / L);
}

// min_image for |d| < L/4 without the division: d/L rounds into (-1/4, 1/4),
// rint gives ±0 and d − L·(±0) = d (up to the sign of a zero d, which a
// squared offset does not see); other offsets take min_image.
__device__ __forceinline__ double min_image_small(double d, double L) {
    return fabs(d) < 0.25 * L ? d : min_image(d, L);
}

// minimum_image of a neighbour offset |d| < L/2 - cutoff or > L/2 + cutoff
// (boxes that are not small): d·(1/L) rounds to the integer d/L rounds to, so
// the result equals min_image's.
__device__ __forceinline__ double min_image_nb(double d, double L, double invL, bool small) {
    return small ? min_image(d, L) : d - L * rint(d * invL);
}

// wrap (src/box.jl:21-30). For 0 < x < L the correctly rounded x/L is at
// most 1 − 2^-53 (the largest double below L is L(1 − δ) with δ ≥ 2^-53), so
// s = 0 and x − 0·L = x: that case returns x without the division
// (bit-identical).
__device__ __forceinline__ double wrap(double x, double L, int32_t &shift) {
#pragma clang fp contract(off)
    if (x > 0.0 && x < L) {
        shift = 0;
        return x;
    }
    double s = floor(x / L);
    double xp = x - s * L;
    if (xp >= L) {
        xp = 0.0;
        s += 1.0;
    }
    shift = (int32_t)s;
    return xp;
}

// ------------------------------------------------------------- body terms

// contour_integrands of the built-ins (src/energies/body_terms.jl:12,33;
// Bending test/test_energies.jl:9).
template <typename T>
__device__ __forceinline__ T term_q(int kind, T R, T dR, T ddR) {
    if (kind == FCC_TERM_BENDING) {
        T s2 = R * R + dR * dR;
        T num = R * R + T(2) * dR * dR - R * ddR;
        T is2 = T(1) / s2;
        return num * num * is2 * is2 / sqrt(s2);
    }
    return sqrt(R * R + dR * dR);
}

// contour_integrand_gradients (src/energies/body_terms.jl:14-17; closed forms
// of the ForwardDiff derivatives of the other two).
template <typename T>
__device__ __forceinline__ void term_J(int kind, T R, T dR, T ddR, T &j0, T &j1, T &j2) {
    if (kind == FCC_TERM_BENDING) {
        T s2 = R * R + dR * dR;
        T num = R * R + T(2) * dR * dR - R * ddR;
        T is2 = T(1) / s2;
        T s5 = is2 * is2 / sqrt(s2);
        T s7 = s5 * is2;
        j0 = T(2) * num * (T(2) * R - ddR) * s5 - T(5) * num * num * R * s7;
        j1 = T(2) * num * (T(4) * dR) * s5 - T(5) * num * num * dR * s7;
        j2 = T(-2) * num * R * s5;
        return;
    }
    T s = sqrt(R * R + dR * dR);
    j0 = R / s;
    j1 = dR / s;
    j2 = T(0);
}

// energy and energy_gradient (src/energies/body_terms.jl:13,18,34).
template <typename T>
__device__ __forceinline__ T term_E(int kind, T p0, T p1, T Q, T r0, T &dF) {
    if (kind == FCC_TERM_PERIMETER_ELASTICITY) {
        T x = Q - p1 * T(1.7724538509055159) * r0;
        dF = p0 * (T(2) * x);
        return p0 * (x * x);
    }
    dF = p0;
    return p0 * Q;
}

// Register layout of a coefficient row with at most KM modes: l[0] = a0,
// l[m] = cosine coefficient of mode m+1, l[KM+m] = sine coefficient (m =
// 1..K; unused slots are zero). Indices stay compile-time for any runtime K.
template <typename T, int KM>
__device__ __forceinline__ void load_row(T (&l)[2 * KM + 1], const T *row, int K) {
    l[0] = row[0];
#pragma unroll
    for (int m = 1; m <= KM; m++) {
        l[m] = m <= K ? row[m] : T(0);
        l[KM + m] = m <= K ? row[K + m] : T(0);
    }
}

template <typename T, int KM>
__device__ __forceinline__ void store_row(const T (&l)[2 * KM + 1], T *row, int K) {
    row[0] = l[0];
#pragma unroll
    for (int m = 1; m <= KM; m++) {
        if (m <= K) {
            row[m] = l[m];
            row[K + m] = l[KM + m];
        }
    }
}

// Rows padded to W elements (a multiple of 16 bytes) are read and written with
// 16-byte vector accesses; every row starts at a multiple of ldc·sizeof(T) ≥ 64
// bytes from a 256-byte aligned allocation, and ldc ≥ W.
template <int NC, typename T>
constexpr int row_width() {
    return ((NC * (int)sizeof(T) + 15) / 16) * 16 / (int)sizeof(T);
}

template <typename T, int W>
__device__ __forceinline__ void row_load(T (&l)[W], const T *row) {
    static_assert((W * sizeof(T)) % 16 == 0, "row width");
    const uint4 *s = reinterpret_cast<const uint4 *>(row);
    uint4 tmp[W * sizeof(T) / 16];
#pragma unroll
    for (int q = 0; q < (int)(W * sizeof(T) / 16); q++) tmp[q] = s[q];
    __builtin_memcpy(l, tmp, sizeof(tmp));
}

template <typename T, int W>
__device__ __forceinline__ void row_store(const T (&l)[W], T *row) {
    static_assert((W * sizeof(T)) % 16 == 0, "row width");
    uint4 tmp[W * sizeof(T) / 16];
    __builtin_memcpy(tmp, l, sizeof(tmp));
    uint4 *s = reinterpret_cast<uint4 *>(row);
#pragma unroll
    for (int q = 0; q < (int)(W * sizeof(T) / 16); q++) s[q] = tmp[q];
}

// contour_radius_dir without the R0 factor (src/geometry.jl:147-156).
template <typename T, int KM>
__device__ __forceinline__ T contour_dir(const T (&l)[2 * KM + 1], int K, T c1, T s1) {
    T c = c1, s = s1, r = l[0];
#pragma unroll
    for (int m = 1; m <= KM; m++) {
        if (m <= K) {
            T cn = c * c1 - s * s1;
            T sn = s * c1 + c * s1;
            c = cn;
            s = sn;
            r += l[m] * c + l[KM + m] * s;
        }
    }
    return r;
}

}  // namespace fcc
