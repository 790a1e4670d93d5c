// fcc_kernels.hpp — gfx950 kernels of the time step (VALU path).
//
// Layout: one work item per body ("body-parallel"), 64-lane workgroups. The
// contour R, R', R'' is never stored in HBM: it is synthesised from the
// body's lab-frame coefficients in registers, node by node, with the basis
// rows read through the scalar cache (uniform node index), and the pullback
// to the coefficients is accumulated in registers in the same pass.
#pragma once

#include "fcc_device.hpp"

namespace fcc {

// Basis row stride on the device: [a0 | cos modes m = 1..31 | sin modes at SH+m].
constexpr int SB = 64;
constexpr int SH = SB / 2;

template <typename T, int KM>
__device__ __forceinline__ void synth(const FCC_CONST T *S, int nth, int k, const T (&l)[2 * KM + 1],
                                      int K, bool need_dd, T r0, T &R, T &dR, T &ddR) {
    const FCC_CONST T *C = S + (size_t)k * SB;
    const FCC_CONST T *D = S + (size_t)(nth + k) * SB;
    const FCC_CONST T *DD = S + (size_t)(2 * nth + k) * SB;
    T r = l[0], dr = T(0), ddr = T(0);
#pragma unroll
    for (int m = 1; m <= KM; m++) {
        if (m <= K) {
            r += C[m] * l[m] + C[SH + m] * l[KM + m];
            dr += D[m] * l[m] + D[SH + m] * l[KM + m];
            if (need_dd) ddr += DD[m] * l[m] + DD[SH + m] * l[KM + m];
        }
    }
    R = r0 * r;
    dR = r0 * dr;
    ddR = r0 * ddr;
}

// Forces: body terms (src/gradients.jl:70-128), Overlap contact
// (src/energies/contact.jl:58-82,149-187) and the pullback to the lab frame
// (src/gradients.jl:374-386), fused. Writes glab (∂E/∂(R0·lab)), gx, gy,
// gθ and Ebody. With use_status it also performs the status mirror and the
// clock tick of a step (src/utils.jl:50-78).
template <typename T, int NTH_, int KM, typename PP, typename DP>
__device__ __forceinline__ void k_forces_impl(const PP &p, const DP &d, int use_status,
                                               int debug, int64_t blk) {
    constexpr int NTHM = NTH_ ? NTH_ : FUSED_MAX_THETA;
    constexpr int NCM = 2 * KM + 1;
    __shared__ T sR[NTHM][64];
    __shared__ T sP[NTHM][64];
    __shared__ T scs[2][NTHM];
    const int nth = NTH_ ? NTH_ : p.nth;
    const int K = p.K;
    const int lane = threadIdx.x;
    for (int k = lane; k < nth; k += 64) {
        scs[0][k] = d.cs[2 * k];
        scs[1][k] = d.cs[2 * k + 1];
    }
    __syncthreads();
    const int64_t i = (int64_t)blk * 64 + lane;
    if (forces_frozen(use_status, d.st, blk, lane)) return;
    if (i >= p.N) return;
    const FCC_CONST T *S = as_const(d.S);
    T lab[NCM];
    load_row<T, KM>(lab, d.lab + i * p.ldc, K);
    const T r0 = d.R0[i];
    const int ti = d.type[i];
    const uint32_t tbit = 1u << ti;
    const bool need_dd = p.need_dd != 0;
    const bool contact = p.ncontact != 0;

    // pass 1: contour integrals Q of every body term (sum over nodes k = 1..nθ)
    T Q[MAXB];
#pragma unroll
    for (int b = 0; b < MAXB; b++) Q[b] = T(0);
    for (int k = 0; k < nth; k++) {
        T R, dR, ddR;
        synth<T, KM>(S, nth, k, lab, K, need_dd, r0, R, dR, ddR);
        if (contact) sR[k][lane] = R;
#pragma unroll
        for (int b = 0; b < MAXB; b++)
            if (b < p.nbody) Q[b] += term_q<T>(p.bkind[b], R, dR, ddR);
    }
    T w[MAXB], Et[MAXB];
#pragma unroll
    for (int b = 0; b < MAXB; b++) {
        w[b] = T(0);
        Et[b] = T(0);
        if (b < p.nbody && (p.bmask[b] & tbit)) {
            T dF;
            Et[b] = term_E<T>(p.bkind[b], p.bp0[b], p.bp1[b], Q[b] * p.dphi, r0, dF);
            w[b] = dF * p.dphi;
        }
    }
    T E = T(0);  // right fold e1 + (e2 + (... + 0)), src/gradients.jl:70-74
#pragma unroll
    for (int b = MAXB - 1; b >= 0; b--)
        if (b < p.nbody) E = Et[b] + E;

    // contact density P at every node, accumulated over the neighbour list in
    // list order; only the nodes inside the angular window that can reach the
    // neighbour's bounding disc are tested.
    if (contact) {
        for (int k = 0; k < nth; k++) sP[k][lane] = T(0);
        const double xi = d.x[i], yi = d.y[i];
        const T rmi = d.rmax[i];
        const int pi = d.parent[i];
        const T Lxf = (T)p.Lx, Lyf = (T)p.Ly;
        int cnt = d.ncount[i];
        if (cnt > d.cap) cnt = d.cap;
        for (int q = 0; q < cnt; q++) {
            const int j = d.nidx[(int64_t)q * p.N + i];
            const T dxij = (T)min_image(xi - d.x[j], p.Lx);
            const T dyij = (T)min_image(yi - d.y[j], p.Ly);
            const T lim = d.rmax[j] + p.range;
            const T lim2 = lim * lim;
            const T D2 = dxij * dxij + dyij * dyij;
            const T reach = (rmi + lim) * T(1.0001) + T(1e-6);
            if (D2 > reach * reach) continue;
            int k0 = 0, wlen = nth;
            const T sd2 = lim2 * T(1.0002) / D2;
            if (sd2 < T(1)) {
                const T delta = asin(sqrt(sd2)) + T(2e-3);
                const T psi = atan2(-dyij, -dxij);
                const int klo = (int)floor((psi - delta) / p.dphi);
                const int khi = (int)ceil((psi + delta) / p.dphi);
                if (khi - klo + 1 < nth) {
                    wlen = khi - klo + 1;
                    k0 = klo % nth;
                    if (k0 < 0) k0 += nth;
                }
            }
            uint64_t mask = 0;
            int k = k0;
            for (int t = 0; t < wlen; t++) {
                const T Rk = sR[k][lane];
                const T dx = min_image(dxij + Rk * scs[0][k], Lxf);
                const T dy = min_image(dyij + Rk * scs[1][k], Lyf);
                if (dx * dx + dy * dy <= lim2) mask |= 1ull << t;
                if (++k == nth) k = 0;
            }
            if (!mask) continue;
            T lj[NCM];
            load_row<T, KM>(lj, d.lab + (int64_t)j * p.ldc, K);
            const T r0j = d.R0[j];
            const T Jv = p.Jc[pi == d.parent[j] ? 1 : 0][ti][d.type[j]];
            const T kr = p.krep[ti];
            while (mask) {
                const int t = __builtin_ctzll(mask);
                mask &= mask - 1;
                int kk = k0 + t;
                if (kk >= nth) kk -= nth;
                const T Rk = sR[kk][lane];
                const T dx = min_image(dxij + Rk * scs[0][kk], Lxf);
                const T dy = min_image(dyij + Rk * scs[1][kk], Lyf);
                const T dist = sqrt(dx * dx + dy * dy);
                const T cp = dist > T(0) ? dx / dist : T(1);
                const T sp = dist > T(0) ? dy / dist : T(0);
                const T dd = r0j * contour_dir<T, KM>(lj, K, cp, sp) - dist;
                sP[kk][lane] += (dd >= T(0) ? Jv : T(0)) + (dd > T(0) ? kr * dd : T(0));
            }
        }
    }

    // pass 2: contour gradients per node, contact force/torque, pullback
    T gl[NCM];
#pragma unroll
    for (int c = 0; c < NCM; c++) gl[c] = T(0);
    T fx = T(0), fy = T(0), tt = T(0);
    for (int k = 0; k < nth; k++) {
        T R, dR, ddR;
        synth<T, KM>(S, nth, k, lab, K, need_dd, r0, R, dR, ddR);
        if (contact) R = sR[k][lane];
        T GR = T(0), GD = T(0), GDD = T(0);
#pragma unroll
        for (int b = 0; b < MAXB; b++) {
            if (b < p.nbody && (p.bmask[b] & tbit)) {
                T j0, j1, j2;
                term_J<T>(p.bkind[b], R, dR, ddR, j0, j1, j2);
                GR += j0 * w[b];
                GD += j1 * w[b];
                GDD += j2 * w[b];
            }
        }
        if (contact) {
            const T g = sP[k][lane];
            const T pp = g * p.dphi;
            const T c = scs[0][k], s = scs[1][k];
            GR += pp * R;
            fx += pp * (R * c + dR * s);
            fy += pp * (R * s - dR * c);
            tt += (-pp) * R * dR;
            if (debug) d.P[i * nth + k] = g;
        }
        const FCC_CONST T *C = S + (size_t)k * SB;
        const FCC_CONST T *D = S + (size_t)(nth + k) * SB;
        const FCC_CONST T *DD = S + (size_t)(2 * nth + k) * SB;
        gl[0] += C[0] * GR;
#pragma unroll
        for (int m = 1; m <= KM; m++) {
            if (m <= K) {
                T a = C[m] * GR + D[m] * GD;
                T b = C[SH + m] * GR + D[SH + m] * GD;
                if (need_dd) {
                    a += DD[m] * GDD;
                    b += DD[SH + m] * GDD;
                }
                gl[m] += a;
                gl[KM + m] += b;
            }
        }
        if (debug) {
            d.GR[i * nth + k] = GR;
            d.GD[i * nth + k] = GD;
            d.GDD[i * nth + k] = GDD;
        }
    }
    store_row<T, KM>(gl, d.glab + i * p.ldc, K);
    d.gx[i] = fx;
    d.gy[i] = fy;
    d.gth[i] = tt;
    d.Ebody[i] = E;
    if (d.th0) d.th0[i] = d.th[i];  // parent drag: θ at the start of the step
}

template <typename T, int NTH_, int KM>
__global__ __launch_bounds__(64) void k_forces(const Params<T> p, const Ptrs<T> d, int use_status,
                                               int debug) {
    k_forces_impl<T, NTH_, KM>(p, d, use_status, debug, blockIdx.x);
}

template <typename T, int NTH_, int KM>
__global__ __launch_bounds__(64) void k_forces_ens(const Params<T> *ps, const Ptrs<T> *ds, const Tile *tiles,
                                                   int use_status, int debug) {
    const Tile tl = load_uniform(tiles, blockIdx.x);
    if (tl.rep < 0) return;
    const FCC_CONST Params<T> &p = as_const(ps)[tl.rep];
    const FCC_CONST Ptrs<T> &d = as_const(ds)[tl.rep];
    k_forces_impl<T, NTH_, KM>(p, d, use_status, debug, tl.blk);
}

// ----------------------------------------------------------- per-body math

// (cos nθ, sin nθ) for n = 2..K+1 in slots 1..K: direct in FP64, by rotation
// from (cos θ, sin θ) in FP32 (src/geometry.jl:132-138 recurrence), where slot 0
// also receives (cos θ, sin θ) (FP64 leaves slot 0 unset).
template <typename T, int KM>
__device__ __forceinline__ void harmonics(T th, int K, T (&cn)[KM + 1], T (&sn)[KM + 1]) {
#pragma clang fp contract(off)
    if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int m = 1; m <= KM; m++)
            if (m <= K) fsincos((T)(m + 1) * th, &sn[m], &cn[m]);
    } else {
        T c1, s1;
        sincosf(th, &s1, &c1);
        cn[0] = c1;
        sn[0] = s1;
        T c = c1, s = s1;
#pragma unroll
        for (int m = 1; m <= KM; m++) {
            T c2 = c * c1 - s * s1;
            T s2 = s * c1 + c * s1;
            c = c2;
            s = s2;
            cn[m] = c;
            sn[m] = s;
        }
    }
}

// enforce_shape! (src/dynamics.jl:44-64) on a register row, in U precision.
// With U = double and T = float, the bound test first takes the sum of the
// mode amplitudes in FP32: each term sqrt(a² + b²) is within 3 ulps of the
// exact one (two squares and their sum, halved by the root, and the
// hardware root's own ≤ 1 ulp; a subnormal argument may flush, an absolute
// error below 2^-63), and summing 15 or fewer non-negative terms adds at
// most 14 roundings, so the FP32 sum is within 17·2^-24 < 2^-19 (relative)
// of the exact one, and the FP64 sum within 2^-50 of it. An FP32 sum below
// bound·(1 − 2^-16) thus decides "no rescale" as the FP64 sum would, and the
// FP64 sum (seven FP64 square roots at K = 7) is taken only near or above
// the bound.
template <typename U, typename T, int KM>
__device__ __forceinline__ void enforce_shape(T (&c)[2 * KM + 1], int K) {
#pragma clang fp contract(off)
    const U bound = U(0.816496580927726);
    bool near = true;
    if constexpr (sizeof(U) == 8 && sizeof(T) == 4 && KM <= 15) {
        float st = 0.f;
#pragma unroll
        for (int m = 1; m <= KM; m++)
            if (m <= K) st += __builtin_amdgcn_sqrtf(c[m] * c[m] + c[KM + m] * c[KM + m]);
        near = !(st < (float)(0.816496580927726 * (1.0 - 1.0 / 65536.0)));
    }
    U sm = U(0);
    if (near) {
#pragma unroll
        for (int m = 1; m <= KM; m++)
            if (m <= K) sm += sqrt((U)c[m] * (U)c[m] + (U)c[KM + m] * (U)c[KM + m]);
    }
    U f = U(1);
    const bool scale = near && sm > bound;
    if (scale) f = bound / sm;
    U q = U(0);
#pragma unroll
    for (int m = 1; m <= KM; m++) {
        if (m <= K) {
            U a = (U)c[m], b = (U)c[KM + m];
            if (scale) {
                a *= f;
                b *= f;
                c[m] = (T)a;
                c[KM + m] = (T)b;
            }
            q += a * a + b * b;
        }
    }
    c[0] = (T)sqrt(U(1) - q / U(2));
}

// Square root of a term of the radius bound: the hardware root in FP32
// (within one ulp of the correctly rounded one, which the FP32 sum of the
// terms rounds as much anyway; rmax and rmin enter the list cutoff, with the
// skin, and the contact tests, next to an FP32-evaluated radius), the
// correctly rounded one in FP64.
__device__ __forceinline__ float bound_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double bound_sqrt(double x) { return sqrt(x); }

// lab_frame! (src/geometry.jl:73-92) on registers.
template <typename T, int KM>
__device__ __forceinline__ T lab_frame(const T (&c)[2 * KM + 1], T th, T r0, int K, T (&l)[2 * KM + 1],
                                       T &rmin) {
#pragma clang fp contract(off)
    T cn[KM + 1], sn[KM + 1];
    harmonics<T, KM>(th, K, cn, sn);
    l[0] = c[0];
    T bound = c[0];
#pragma unroll
    for (int m = 1; m <= KM; m++) {
        l[m] = T(0);
        l[KM + m] = T(0);
        if (m <= K) {
            T a = c[m], b = c[KM + m];
            l[m] = a * cn[m] - b * sn[m];
            l[KM + m] = a * sn[m] + b * cn[m];
            bound += bound_sqrt(a * a + b * b);
        }
    }
    rmin = r0 * (c[0] - (bound - c[0]));
    return r0 * bound;
}

// Inverse of lab_frame on the coefficient rows: the body-frame coefficients
// (a_n, b_n) from the lab-frame row l (rotation by -nθ). The lab rows are the
// stored state; the body-frame coefficients are derived from them where they
// are needed (the update, body_frame!, state reads).
template <typename T, int KM>
__device__ __forceinline__ void lab_to_coef(T (&c)[2 * KM + 1], const T (&cn)[KM + 1], const T (&sn)[KM + 1],
                                            int K) {
#pragma clang fp contract(off)
#pragma unroll
    for (int m = 1; m <= KM; m++) {
        if (m <= K) {
            const T al = c[m], bl = c[KM + m];
            c[m] = al * cn[m] + bl * sn[m];
            c[KM + m] = bl * cn[m] - al * sn[m];
        }
    }
}

template <typename T, int KM>
__device__ __forceinline__ void lab_to_coef(T (&c)[2 * KM + 1], T th, int K) {
    T cn[KM + 1], sn[KM + 1];
    harmonics<T, KM>(th, K, cn, sn);
    lab_to_coef<T, KM>(c, cn, sn, K);
}

// Construction / refresh: enforce_shape! then lab_frame! (src/simulation.jl:102-103).
template <typename T, int KM, bool F64S>
__global__ __launch_bounds__(256) void k_init(const Params<T> p, const Ptrs<T> d) {
    constexpr int NCM = 2 * KM + 1;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= p.N) return;
    T c[NCM], l[NCM];
    load_row<T, KM>(c, d.coef + i * p.ldc, p.K);
    if constexpr (F64S)
        enforce_shape<double, T, KM>(c, p.K);
    else
        enforce_shape<T, T, KM>(c, p.K);
    store_row<T, KM>(c, d.coef + i * p.ldc, p.K);
    T rmn;
    d.rmax[i] = lab_frame<T, KM>(c, d.th[i], d.R0[i], p.K, l, rmn);
    d.rmin[i] = rmn;
    store_row<T, KM>(l, d.lab + i * p.ldc, p.K);
}

// Quotients x / d of many numerators by one divisor, each correctly rounded
// (the bits of `x / d`). FP64: the division itself. FP32: one binary64
// reciprocal rd = RN(1/d) shared by all numerators, then RN32(x·rd). The
// product is within 2^-51.9 (relative) of q = x/d. A quotient of two 24-bit
// significands is never a binary32 rounding midpoint m, and q − m =
// (x − m·d)/d for q in [2^E, 2^(E+1)) and d in [2^e_d, 2^(e_d+1)): m is an
// odd multiple of 2^(E−24), d a multiple of 2^(e_d−23) and x (≥ 2^(E+e_d)) a
// multiple of 2^(E+e_d−23), so x − m·d is a nonzero multiple of
// 2^(E+e_d−47) and |q − m| ≥ 2^(E+e_d−47) / 2^(e_d+1) = 2^(E−48), which is
// more than the product's error (< 2^(E−50.9)): no midpoint lies between q and
// the product, and both round to the same binary32 value (the subnormal range
// keeps a margin of the same kind; zero, infinite and NaN operands give the
// IEEE results).
template <typename T>
struct SharedDivisor {
    T d;
    __device__ __forceinline__ explicit SharedDivisor(T d_) : d(d_) {}
    __device__ __forceinline__ T div(T x) const { return x / d; }
};
template <>
struct SharedDivisor<float> {
    double rd;
    __device__ __forceinline__ explicit SharedDivisor(float d) : rd(1.0 / (double)d) {}
    __device__ __forceinline__ float div(float x) const { return (float)((double)x * rd); }
};

// body_frame! (src/gradients.jl:347-362): gcoef from glab at the current
// state, given the harmonics of θ.
template <typename T, int KM>
__device__ __forceinline__ void body_frame(const T (&gl)[2 * KM + 1], const T (&c)[2 * KM + 1],
                                           const T (&cn)[KM + 1], const T (&sn)[KM + 1], T r0, int K,
                                           T (&gc)[2 * KM + 1]) {
#pragma clang fp contract(off)
    const T ga0 = r0 * gl[0];
    const T a0 = c[0];
    const SharedDivisor<T> dv(T(2) * a0);
    gc[0] = ga0;
#pragma unroll
    for (int m = 1; m <= KM; m++) {
        gc[m] = T(0);
        gc[KM + m] = T(0);
        if (m <= K) {
            T ga = r0 * gl[m], gb = r0 * gl[KM + m];
            T a = c[m], b = c[KM + m];
            gc[m] = ga * cn[m] + gb * sn[m] - dv.div(ga0 * a);
            gc[KM + m] = (-ga) * sn[m] + gb * cn[m] - dv.div(ga0 * b);
        }
    }
}

template <typename T, int KM>
__global__ __launch_bounds__(256) void k_body_frame(const Params<T> p, const Ptrs<T> d) {
    constexpr int NCM = 2 * KM + 1;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= p.N) return;
    T gl[NCM], c[NCM], gc[NCM];
    const T th = d.th[i];
    load_row<T, KM>(gl, d.glab + i * p.ldc, p.K);
    load_row<T, KM>(c, d.lab + i * p.ldc, p.K);
    T cn[KM + 1], sn[KM + 1];
    harmonics<T, KM>(th, p.K, cn, sn);
    lab_to_coef<T, KM>(c, cn, sn, p.K);
    body_frame<T, KM>(gl, c, cn, sn, d.R0[i], p.K, gc);
    store_row<T, KM>(gc, d.gcoef + i * p.ldc, p.K);
}

// Body-frame coefficients of the current state into `coef` (before state
// reads and checkpoints; the time step itself keeps only the lab rows).
template <typename T, int KM>
__global__ __launch_bounds__(256) void k_coef_from_lab(const Params<T> p, const Ptrs<T> d) {
    constexpr int NCM = 2 * KM + 1;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= p.N) return;
    T c[NCM];
    load_row<T, KM>(c, d.lab + i * p.ldc, p.K);
    lab_to_coef<T, KM>(c, d.th[i], p.K);
    store_row<T, KM>(c, d.coef + i * p.ldc, p.K);
}

// substrate_displacement (src/dynamics.jl:84-99) of body c with its
// orientation thc.
template <typename T, typename PP, typename DP>
__device__ __forceinline__ void substrate_disp(const PP &p, const DP &d, int64_t c, T thc, uint64_t draw, T &ux,
                                               T &uy) {
    const int tc = d.type[c];
    T fx = d.gx[c], fy = d.gy[c];
    const int nch = d.ccount[c];
    for (int q = 0; q < nch; q++) {
        const int k = d.cidx[(int64_t)q * p.N + c];
        if (p.pdrag[d.type[k]]) {
            fx += d.gx[k];
            fy += d.gy[k];
        }
    }
    T s, co;
    fsincos(thc, &s, &co);
    const T st = p.sDt[tc];
    T x1 = T(0), x2 = T(0);
    if (st != T(0)) philox_pair<T>(p.key0, p.key1, (uint32_t)(c + 1), draw, 0, x1, x2);
    ux = p.dt * (p.v0[tc] * co - p.mur[tc] * fx) + st * x1;
    uy = p.dt * (p.v0[tc] * s - p.mur[tc] * fy) + st * x2;
}

// displacement_kernel! (src/dynamics.jl:128-147) of body i (type ti,
// orientation thi, centre force (gx, gy)) when some body type has parent
// drag: a child adds its parent's substrate displacement, whose orientation
// is the parent's at the start of the step (d.th0, stored by the forces
// kernel); a parent's includes the forces on its children with parent drag.
// d2 is the rotational noise increment.
template <typename T, typename PP, typename DP>
__device__ __forceinline__ void pdrag_disp(const PP &p, const DP &d, int64_t i, int ti, T thi, T gx, T gy,
                                           uint64_t draw, T &d0, T &d1, T &d2) {
    if (p.pdrag[ti]) {
        const int64_t pr = d.parent[i];
        T ux, uy;
        substrate_disp<T>(p, d, pr, d.th0[pr], draw, ux, uy);
        T s, co;
        fsincos(thi, &s, &co);
        const T st = p.sDt[ti];
        T x1 = T(0), x2 = T(0);
        if (st != T(0)) philox_pair<T>(p.key0, p.key1, (uint32_t)(i + 1), draw, 0, x1, x2);
        d0 = ux + p.dt * (p.v0[ti] * co - p.mur[ti] * gx) + st * x1;
        d1 = uy + p.dt * (p.v0[ti] * s - p.mur[ti] * gy) + st * x2;
    } else {
        substrate_disp<T>(p, d, i, thi, draw, d0, d1);
    }
    T x3 = T(0), x4;
    if (p.sDr[ti] != T(0)) philox_pair<T>(p.key0, p.key1, (uint32_t)(i + 1), draw, 1, x3, x4);
    d2 = p.sDr[ti] * x3;
}

// Pieces of the Euler–Maruyama step. They and the helpers they call
// (harmonics, body_frame, enforce_shape, lab_frame, wrap, min_image) are
// compiled without FMA contraction, as the reference's arithmetic is (Julia
// does not contract): the same operations give the same bits wherever they
// are inlined.

// Parameters of one body type used by the update, taken from the uniform
// parameter block by a select chain over the types (no per-lane load that
// would wait on the body's type).
template <typename T>
struct TypeParams {
    T v0, mur, sDt, sDr, muth, muab;  // sDt, sDr: √(2·Dt·dt), √(2·Dr·dt)
};

// (readfirstlane keeps the compiler from folding the chain back into a load
// indexed by the type.)
template <typename T>
__device__ __forceinline__ T opaque_uniform(T v) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    }
}

template <typename T, typename A>
__device__ __forceinline__ T type_param(const A &a, int ti) {
    T at[FCC_MAX_TYPES];
#pragma unroll
    for (int t = 0; t < FCC_MAX_TYPES; t++) at[t] = opaque_uniform<T>(a[t]);
    T v = at[0];
#pragma unroll
    for (int t = 1; t < FCC_MAX_TYPES; t++) v = ti == t ? at[t] : v;
    return v;
}

template <typename T, typename PP>
__device__ __forceinline__ TypeParams<T> type_params(const PP &p, int ti) {
    TypeParams<T> tp;
    if (p.one_type) {  // uniform: every body is of type 0
        tp.v0 = p.v0[0];
        tp.mur = p.mur[0];
        tp.sDt = p.sDt[0];
        tp.sDr = p.sDr[0];
        tp.muth = p.muth[0];
        tp.muab = p.muab[0];
        return tp;
    }
    tp.v0 = type_param<T>(p.v0, ti);
    tp.mur = type_param<T>(p.mur, ti);
    tp.sDt = type_param<T>(p.sDt, ti);
    tp.sDr = type_param<T>(p.sDr, ti);
    tp.muth = type_param<T>(p.muth, ti);
    tp.muab = type_param<T>(p.muab, ti);
    return tp;
}

// Displacement of a body without parent drag (src/dynamics.jl:84-99):
// self-propulsion along θ (s, co = sin θ, cos θ), mobility times the centre
// force, translational noise from Philox block 0 of draw `draw`
// (src/dynamics.jl:128-147).
template <typename T, typename PP>
__device__ __forceinline__ void body_displacement(const PP &p, int64_t i, const TypeParams<T> &tp, T s, T co,
                                                  T gx, T gy, uint64_t draw, T &d0, T &d1) {
#pragma clang fp contract(off)
    T x1 = T(0), x2 = T(0);
    if (tp.sDt != T(0)) philox_pair<T>(p.key0, p.key1, (uint32_t)(i + 1), draw, 0, x1, x2);
    const T st = tp.sDt;
    d0 = p.dt * (tp.v0 * co - tp.mur * gx) + st * x1;
    d1 = p.dt * (tp.v0 * s - tp.mur * gy) + st * x2;
}

// Rotational noise increment (Philox block 1 of draw `draw`).
template <typename T, typename PP>
__device__ __forceinline__ T rotation_noise(const PP &p, int64_t i, const TypeParams<T> &tp, uint64_t draw) {
#pragma clang fp contract(off)
    T x3 = T(0), x4;
    if (tp.sDr != T(0)) philox_pair<T>(p.key0, p.key1, (uint32_t)(i + 1), draw, 1, x3, x4);
    return tp.sDr * x3;
}

// Body-frame gradient, orientation and shape update with the shape
// constraint (src/dynamics.jl:180-203, src/gradients.jl:347-362; the
// constraint in FP64 when F64S), given the harmonics of θ. Returns the new θ;
// c holds the new coefficients and gc the body-frame gradient.
template <typename T, int KM, bool F64S, typename PP>
__device__ __forceinline__ T shape_step(const PP &p, const TypeParams<T> &tp, int K, const T (&gl)[2 * KM + 1],
                                        T (&c)[2 * KM + 1], const T (&cn)[KM + 1], const T (&sn)[KM + 1], T th,
                                        T r0, T gth, T d2, T (&gc)[2 * KM + 1]) {
#pragma clang fp contract(off)
    body_frame<T, KM>(gl, c, cn, sn, r0, K, gc);
    const T thn = th - p.dt * tp.muth * gth + d2;
    const T h = p.dt * tp.muab;
#pragma unroll
    for (int m = 1; m <= KM; m++) {
        c[m] -= h * gc[m];
        c[KM + m] -= h * gc[KM + m];
    }
    if constexpr (F64S)
        enforce_shape<double, T, KM>(c, K);
    else
        enforce_shape<T, T, KM>(c, K);
    return thn;
}

// Position wrap in FP64 (src/box.jl:21-30), the image counters, the lab
// frame of the new state (src/geometry.jl:73-92) and the Verlet skin check
// (src/neighbors.jl:198-205) of body i, given its position (x, y), its new
// coefficients c and orientation thn, its displacement (d0, d1) and the
// skin reference (x0, y0, rmax0). Returns whether a body of the wave tripped
// the skin check; `trip` (if not null) is then set to 1.
template <typename T, int KM, bool EXACT, typename PP, typename DP>
__device__ __forceinline__ bool place_body(const PP &p, const DP &d, int64_t i, int K, const T (&c)[2 * KM + 1],
                                           T thn, T r0, T d0, T d1, double x, double y, double x0, double y0,
                                           T rmax0, int32_t *trip) {
#pragma clang fp contract(off)
    constexpr int NCM = 2 * KM + 1;
    constexpr int W = row_width<NCM, T>();
    T l[NCM];
    int32_t sx, sy;
    const double xw = wrap(x + (double)d0, p.Lx, sx);
    const double yw = wrap(y + (double)d1, p.Ly, sy);
    T rmn;
    const T rm = lab_frame<T, KM>(c, thn, r0, K, l, rmn);
    d.x[i] = xw;
    d.y[i] = yw;
    if (sx != 0) d.imgx[i] += sx;  // image counters change only on a wrap (rare): no traffic otherwise
    if (sy != 0) d.imgy[i] += sy;
    d.rmax[i] = rm;
    d.rmin[i] = rmn;
    if constexpr (EXACT) {
        T b[W];
#pragma unroll
        for (int q = 0; q < W; q++) b[q] = q < NCM ? l[q] : T(0);
        row_store<T, W>(b, d.lab + i * p.ldc);
    } else {
        store_row<T, KM>(l, d.lab + i * p.ldc, K);
    }
    if (p.skin_check) {
        const double dx = min_image_small(xw - x0, p.Lx);  // only dx², dy² are used
        const double dy = min_image_small(yw - y0, p.Ly);
        const double dr = (double)rm - (double)rmax0;
        const double delta = sqrt(dx * dx + dy * dy) + (dr > 0.0 ? dr : 0.0);
        // one store per wave whose bodies tripped (not one per body)
        const uint64_t trips = __ballot(2.0 * delta > (double)p.skin);
        if (trip && trips != 0 && (int)(threadIdx.x & 63) == __ffsll((long long)trips) - 1)
            __hip_atomic_store(trip, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return trips != 0;
    }
    return false;
}

template <typename T, int KM, bool EXACT>
__device__ __forceinline__ void store_coef(T *row, const T (&c)[2 * KM + 1], int K) {
    constexpr int NCM = 2 * KM + 1;
    constexpr int W = row_width<NCM, T>();
    if constexpr (EXACT) {
        T a[W];
#pragma unroll
        for (int q = 0; q < W; q++) a[q] = q < NCM ? c[q] : T(0);
        row_store<T, W>(a, row);
    } else {
        store_row<T, KM>(c, row, K);
    }
}

template <typename T, int KM, bool EXACT>
__device__ __forceinline__ void load_coef(T (&c)[2 * KM + 1], const T *row, int K) {
    constexpr int NCM = 2 * KM + 1;
    constexpr int W = row_width<NCM, T>();
    if constexpr (EXACT) {
        T a[W];
        row_load<T, W>(a, row);
#pragma unroll
        for (int q = 0; q < NCM; q++) c[q] = a[q];
    } else {
        load_row<T, KM>(c, row, K);
    }
}

// Per-body inputs of the update.
template <typename T, int KM>
struct BodyIn {
    int ti;
    T th, r0, gth;
    T f0, f1;  // centre force (gx, gy)
    double x, y, x0, y0;
    T rmax0;
    T gl[2 * KM + 1], c[2 * KM + 1];
};

// Issues every load of body i, before any branch on loaded data: one memory
// round trip covers the body's whole state (behind the noise draws, which
// branch on the type's diffusion constants, each load would wait apart).
template <typename T, int KM, bool EXACT, typename PP, typename DP>
__device__ __forceinline__ void load_body(const PP &p, const DP &d, int64_t i, int K, BodyIn<T, KM> &b) {
    b.ti = d.type[i];
    b.th = d.th[i];
    b.r0 = d.R0[i];
    b.gth = d.gth[i];
    b.x = d.x[i];
    b.y = d.y[i];
    b.f0 = d.gx[i];
    b.f1 = d.gy[i];
    load_coef<T, KM, EXACT>(b.gl, d.glab + i * p.ldc, K);
    load_coef<T, KM, EXACT>(b.c, d.lab + i * p.ldc, K);
    b.x0 = d.x0[i];  // the skin reference is allocated with the neighbour list
    b.y0 = d.y0[i];
    b.rmax0 = d.rmax0[i];
}

// The update of body i from its loaded inputs: displacement and noise, θ and
// shape step, wrap, lab frame, skin check and stores. sin θ, cos θ and the
// harmonics of θ are computed once for the displacement, the body-frame
// coefficients and the body-frame gradient. With `pdrag` (some body type has
// parent drag) the displacement is that of displacement_kernel!
// (pdrag_disp). Returns whether a body of the
// wave tripped the skin check (which also sets *trip unless it is null);
// with `budget` body 0 takes the step budget into `stop`.
template <typename T, int KM, bool EXACT, bool F64S, typename PP, typename DP>
__device__ __forceinline__ bool finish_body(const PP &p, const DP &d, int64_t i, int K, int pdrag, int debug,
                                            uint64_t draw, BodyIn<T, KM> &b, int32_t *trip, bool budget) {
    constexpr int NCM = 2 * KM + 1;
    const TypeParams<T> tp = type_params<T>(p, b.ti);
    T cn[KM + 1], sn[KM + 1];
    harmonics<T, KM>(b.th, K, cn, sn);
    T d0, d1, d2;
    if (pdrag) {
        pdrag_disp<T>(p, d, i, b.ti, b.th, b.f0, b.f1, draw, d0, d1, d2);
    } else {
        T s, co;
        if constexpr (sizeof(T) == 4) {
            s = sn[0];
            co = cn[0];
        } else {
            fsincos(b.th, &s, &co);
        }
        body_displacement<T>(p, i, tp, s, co, b.f0, b.f1, draw, d0, d1);
        d2 = rotation_noise<T>(p, i, tp, draw);
    }
    T gc[NCM];
    lab_to_coef<T, KM>(b.c, cn, sn, K);
    const T thn = shape_step<T, KM, F64S>(p, tp, K, b.gl, b.c, cn, sn, b.th, b.r0, b.gth, d2, gc);
    d.th[i] = thn;
    if (debug) store_coef<T, KM, EXACT>(d.gcoef + i * p.ldc, gc, K);
    if (budget && i == 0) d.st->stop = d.st->clock >= d.st->target ? 1 : 0;
    return place_body<T, KM, EXACT>(p, d, i, K, b.c, thn, b.r0, d0, d1, b.x, b.y, b.x0, b.y0, b.rmax0, trip);
}

// Fused Euler–Maruyama update (src/dynamics.jl:180-203) with the
// displacement and Philox noise of src/dynamics.jl:84-99,128-147 (with
// `pdrag`, the displacement_kernel! form with parent drag), the body-frame gradient, the shape constraint (FP64
// when F64S), position wrapping in FP64, the lab frame of the new state and
// the Verlet skin check (src/neighbors.jl:198-205); one body per lane.
template <typename T, int KM, bool EXACT, bool F64S, int UB, typename PP, typename DP>
__device__ __forceinline__ void k_update_impl(const PP &p, const DP &d, int pdrag, int debug, int64_t blk) {
    const int64_t i = (int64_t)blk * (UB ? UB : (int)blockDim.x) + threadIdx.x;
    if (i >= p.N) return;
    const int K = EXACT ? KM : p.K;
    BodyIn<T, KM> b;
    load_body<T, KM, EXACT>(p, d, i, K, b);
    if (d.st->update != 0) return;
    const uint64_t draw = (uint64_t)(d.st->clock - 1);
    (void)finish_body<T, KM, EXACT, F64S>(p, d, i, K, pdrag, debug, draw, b, &d.st->rebuild, true);
}

template <typename T, int KM, bool EXACT, bool F64S>
__global__ __launch_bounds__(256) void k_update(const Params<T> p, const Ptrs<T> d, int pdrag, int debug) {
    k_update_impl<T, KM, EXACT, F64S, 256>(p, d, pdrag, debug, blockIdx.x);
}

// Ensemble launch: any workgroup size up to 256 (the tile size of the table).
template <typename T, int KM, bool EXACT, bool F64S>
__global__ __launch_bounds__(256) void k_update_ens(const Params<T> *ps, const Ptrs<T> *ds, const Tile *tiles,
                                                    int pdrag, int debug) {
    const Tile tl = load_uniform(tiles, blockIdx.x);
    if (tl.rep < 0) return;
    const FCC_CONST Params<T> &p = as_const(ps)[tl.rep];
    const FCC_CONST Ptrs<T> &d = as_const(ds)[tl.rep];
    k_update_impl<T, KM, EXACT, F64S, 0>(p, d, pdrag, debug, tl.blk);
}

// Update part of a one-kernel step (single simulations without parent drag,
// bonds or debug buffers) reading the buffers of parity q at clock clk (the
// status prologue, step_open): the forces part of the workgroup left, for each
// of its 64 bodies, the pullback row (W values, at rows + (l / nb)·stride +
// (l % nb)·W for body l of the workgroup) and the centre force and torque
// (after the nb rows of each slice). One wave updates the 64 bodies, one
// body per lane, with the arithmetic of k_update: it reads this step's state
// and writes positions, lab rows and radius bounds into the next-state
// buffers (d.xn, ...), since other workgroups may still read this step's
// state as neighbours; θ and the image counters stay in place (only the body
// itself reads them). The draw index is the step's clock. A skin trip marks
// tstep[1 - q] with the clock, which freezes the next step; in an ensemble
// step (`gated`: k_gate took the freeze decision into `update` and ticked
// the clock before this launch) it raises the rebuild flag, which no
// workgroup of the launch reads, and body 0 takes the step budget, as in
// k_update. The rows are LDS written before the workgroup barrier that
// precedes this call.
template <typename T, int KM, bool F64S, typename PP, typename DP>
__device__ __forceinline__ void fused_update(const PP &p, const DP &d, const T *rows, int stride, int nb,
                                             int64_t body0, int lane, int q, int64_t clk, bool gated) {
    constexpr int NCM = 2 * KM + 1;
    constexpr int W = row_width<NCM, T>();
    const int64_t i = body0 + lane;
    bool tripped = false;
    if (i < p.N) {
        const T *row = rows + (lane / nb) * stride + (lane % nb) * W;
        const T *f = rows + (lane / nb) * stride + nb * W + 3 * (lane % nb);
        BodyIn<T, KM> b;
        b.ti = d.type[i];
        b.th = d.th[i];
        b.r0 = d.R0[i];
        b.x = d.x[i];
        b.y = d.y[i];
        load_coef<T, KM, true>(b.c, d.lab + i * p.ldc, KM);
        b.x0 = d.x0[i];
        b.y0 = d.y0[i];
        b.rmax0 = d.rmax0[i];
#pragma unroll
        for (int c = 0; c < NCM; c++) b.gl[c] = row[c];
        b.f0 = f[0];
        b.f1 = f[1];
        b.gth = f[2];
        DP dn = d;
        dn.x = d.xn;
        dn.y = d.yn;
        dn.lab = d.labn;
        dn.rmax = d.rmaxn;
        dn.rmin = d.rminn;
        tripped = finish_body<T, KM, true, F64S>(p, dn, i, KM, 0, 0, (uint64_t)clk, b,
                                                  gated ? &d.st->rebuild : nullptr, gated);
    }
    if (!gated && __ballot(tripped) != 0 && lane == 0)
        __hip_atomic_store(&d.st->tstep[q ^ 1], clk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------- cell / neighbours
// No FMA contraction in this section: the decisions (cell index, distance
// test) must equal the reference's IEEE evaluation bit for bit.

// cell_index_kernel! (src/neighbors.jl:41-55) plus the per-cell histogram of
// the counting sort.
__global__ __launch_bounds__(256) void k_cell_index(int64_t N, const double *x, const double *y,
                                                    int32_t ncx, int32_t ncy, double wx, double wy,
                                                    int32_t *cellid, int32_t *count) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    int32_t cx = (int32_t)(x[i] / wx), cy = (int32_t)(y[i] / wy);
    cx = cx < ncx - 1 ? cx : ncx - 1;
    cy = cy < ncy - 1 ? cy : ncy - 1;
    cx = cx > 0 ? cx : 0;  // bodies outside [0, L): clamp (no-op inside the box)
    cy = cy > 0 ? cy : 0;
    const int32_t c = cx + ncx * cy;
    cellid[i] = c;
    atomicAdd(&count[c], 1);
}

// Exclusive scan of the cell counts, start[c] = sum of count[c'] for c' < c,
// in two launches over tiles of SCAN_TILE cells: k_scan_tiles scans each
// tile (coalesced loads, wave prefix sums, one LDS pass over the 16 wave
// totals) and stores the tile total; k_scan_offsets adds the totals of the
// tiles before it (a parallel reduction per tile) and writes start and fill.
constexpr int SCAN_TILE = 1024;

__device__ __forceinline__ int32_t wave_prefix_incl(int32_t v, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t u = __shfl_up(v, off);
        v += lane >= off ? u : 0;
    }
    return v;
}

__global__ __launch_bounds__(SCAN_TILE) void k_scan_tiles(int64_t n, const int32_t *count, int32_t *start,
                                                          int32_t *tile_sum) {
    __shared__ int32_t wsum[SCAN_TILE / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int64_t c = (int64_t)blockIdx.x * SCAN_TILE + t;
    const int32_t v = c < n ? count[c] : 0;
    const int32_t incl = wave_prefix_incl(v, lane);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    if (t < 64) {
        const int32_t x = t < SCAN_TILE / 64 ? wsum[t] : 0;
        const int32_t xi = wave_prefix_incl(x, t);
        if (t < SCAN_TILE / 64) wsum[t] = xi - x;  // exclusive over the waves
    }
    __syncthreads();
    if (c < n) start[c] = wsum[w] + incl - v;
    if (t == SCAN_TILE - 1) tile_sum[blockIdx.x] = wsum[w] + incl;
}

__global__ __launch_bounds__(SCAN_TILE) void k_scan_offsets(int64_t n, const int32_t *tile_sum, int32_t *start,
                                                            int32_t *fill) {
    __shared__ int32_t wsum[SCAN_TILE / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int32_t a = 0;
    for (int64_t b = t; b < (int64_t)blockIdx.x; b += SCAN_TILE) a += tile_sum[b];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) a += __shfl_xor(a, off);
    if (lane == 0) wsum[w] = a;
    __syncthreads();
    int32_t off = 0;
#pragma unroll
    for (int q = 0; q < SCAN_TILE / 64; q++) off += wsum[q];
    const int64_t c = (int64_t)blockIdx.x * SCAN_TILE + t;
    if (c < n) {
        const int32_t s = start[c] + off;
        start[c] = s;
        fill[c] = s;
    }
}

__global__ __launch_bounds__(256) void k_scatter(int64_t N, const int32_t *cellid, int32_t *fill,
                                                 int32_t *perm) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    perm[atomicAdd(&fill[cellid[i]], 1)] = (int32_t)i;
}

// Order each cell's bodies by index: the counting sort becomes the stable
// sortperm! of src/neighbors.jl:86. The positions and bounds of the sorted
// bodies are copied to xs, ys, rs (cell-sorted order) for the neighbour
// kernel, whose candidates are then contiguous runs of these arrays.
template <typename T>
__global__ __launch_bounds__(256) void k_cell_sort(int64_t nce, const int32_t *start, const int32_t *count,
                                                   int32_t *perm, const double *x, const double *y,
                                                   const T *rmax, double *xs, double *ys, T *rs) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= nce) return;
    const int32_t s0 = start[c];
    int32_t *seg = perm + s0;
    const int n = count[c];
    for (int a = 1; a < n; a++) {
        int32_t v = seg[a];
        int b = a - 1;
        while (b >= 0 && seg[b] > v) {
            seg[b + 1] = seg[b];
            b--;
        }
        seg[b + 1] = v;
    }
    for (int a = 0; a < n; a++) {
        const int32_t j = seg[a];
        xs[s0 + a] = x[j];
        ys[s0 + a] = y[j];
        rs[s0 + a] = rmax[j];
    }
}

// neighbor_kernel! (src/neighbors.jl:148-174) for body i: 3×3 cells, oy
// outer, ox inner, cell members in sorted order; entries beyond `cap` are
// counted only. Returns the count. Candidates are read from the cell-sorted
// copies (xs, ys, rs at sorted position q hold the values of body perm[q] bit
// for bit): the cells ox = -1, 0, +1 of a row that does not wrap are
// consecutive cell ids, so their members are one run of sorted positions in
// exactly the reference's order; a wrapping row is scanned cell by cell. A
// run is read in batches of four candidates (position, bound and body id)
// whose loads are issued together.
template <typename T>
__device__ __forceinline__ void neighbor_run(int64_t i, int64_t N, double xi, double yi, double ri, int32_t a,
                                             int32_t b, const int32_t *perm, const double *xs, const double *ys,
                                             const T *rs, double Lx, double Ly, double qLx, double qLy,
                                             int32_t cap, int32_t *nidx, int32_t &n) {
#pragma clang fp contract(off)
    for (int32_t q = a; q < b; q += 4) {
        double cx[4], cy[4];
        T cr[4];
        int32_t cj[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int32_t qq = q + u < b ? q + u : b - 1;  // in bounds: a <= q < b
            cx[u] = xs[qq];
            cy[u] = ys[qq];
            cr[u] = rs[qq];
            cj[u] = perm[qq];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (q + u >= b) break;
            // minimum_image (src/box.jl:14); below a quarter box the rounded
            // quotient is +-0 and the result is dx itself, so the division
            // is skipped there (same value, bit for bit)
            double dx = xi - cx[u];
            if (!(fabs(dx) < qLx)) dx = dx - Lx * rint(dx / Lx);
            double dy = yi - cy[u];
            if (!(fabs(dy) < qLy)) dy = dy - Ly * rint(dy / Ly);
            const double lim = ri + (double)cr[u];
            if (dx * dx + dy * dy <= lim * lim) {
                const int32_t j = cj[u];
                if (j == i) continue;
                n++;
                if (n <= cap) nidx[(int64_t)(n - 1) * N + i] = j;
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ int32_t neighbors_of(int64_t i, int64_t N, double xi, double yi, T rmaxi,
                                                const int32_t *perm, const int32_t *start,
                                                const int32_t *count, const double *xs, const double *ys,
                                                const T *rs, double Lx, double Ly, int32_t ncx,
                                                int32_t ncy, double wx, double wy, double range, int32_t cap,
                                                int32_t *nidx) {
#pragma clang fp contract(off)
    const double ri = (double)rmaxi + range;
    int32_t cx = (int32_t)(xi / wx), cy = (int32_t)(yi / wy);
    cx = cx < ncx - 1 ? cx : ncx - 1;
    cy = cy < ncy - 1 ? cy : ncy - 1;
    cx = cx > 0 ? cx : 0;  // bodies outside [0, L): clamp (no-op inside the box)
    cy = cy > 0 ? cy : 0;
    const int32_t rx = ncx >= 3 ? 1 : 0, ry = ncy >= 3 ? 1 : 0;
    const double qLx = 0.25 * Lx, qLy = 0.25 * Ly;
    // the row's cells are consecutive ids when cx - rx .. cx + rx stay inside
    const bool row_run = cx - rx >= 0 && cx + rx <= ncx - 1;
    int32_t n = 0;
    for (int32_t oy = -ry; oy <= ry; oy++) {
        int32_t yy = (cy + oy) % ncy;
        yy = yy < 0 ? yy + ncy : yy;
        if (row_run) {
            const int32_t c0 = cx - rx + ncx * yy, c1 = cx + rx + ncx * yy;
            neighbor_run<T>(i, N, xi, yi, ri, start[c0], start[c1] + count[c1], perm, xs, ys, rs, Lx, Ly, qLx,
                            qLy, cap, nidx, n);
            continue;
        }
        for (int32_t ox = -rx; ox <= rx; ox++) {
            int32_t xx = (cx + ox) % ncx;
            xx = xx < 0 ? xx + ncx : xx;
            const int32_t cell = xx + ncx * yy;
            const int32_t s0 = start[cell];
            neighbor_run<T>(i, N, xi, yi, ri, s0, s0 + count[cell], perm, xs, ys, rs, Lx, Ly, qLx, qLy, cap,
                            nidx, n);
        }
    }
    return n;
}

// The largest count goes to st->overflow (one atomic per wave). With x0
// non-null the body's position and bound are also stored as the reference
// state of the Verlet skin (src/neighbors.jl:234-238). Workgroups take their
// bodies in XCD-grouped order: each XCD's L2 then holds the candidates of a
// contiguous band of bodies.
template <typename T>
__global__ __launch_bounds__(256) void k_neighbors(int64_t N, const double *x, const double *y, const T *rmax,
                                                   const int32_t *perm, const int32_t *start,
                                                   const int32_t *count, const double *xs, const double *ys,
                                                   const T *rs, double Lx, double Ly, int32_t ncx,
                                                   int32_t ncy, double wx, double wy, double range,
                                                   int32_t cap, int32_t *nidx, int32_t *ncount,
                                                   DevStatus *st, double *x0, double *y0, T *rmax0) {
    const int64_t i = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    int32_t n = 0;
    if (i < N) {
        const double xi = x[i], yi = y[i];
        const T rmi = rmax[i];
        n = neighbors_of<T>(i, N, xi, yi, rmi, perm, start, count, xs, ys, rs, Lx, Ly, ncx, ncy, wx, wy, range,
                            cap, nidx);
        ncount[i] = n;
        if (x0) {
            x0[i] = xi;
            y0[i] = yi;
            rmax0[i] = rmi;
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) n = max(n, __shfl_xor(n, off));
    if ((threadIdx.x & 63) == 0) atomicMax(&st->overflow, n);
}

__global__ __launch_bounds__(256) void k_fill_i32(int64_t n, int32_t *a, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_clear_status(DevStatus *st) {
    st->rebuild = 0;
    st->update = 0;
    st->overflow = 0;
    st->tstep[0] = st->tstep[1] = INT64_MIN;
}

// Before one-kernel steps reading the buffers of parity q: a skin trip of the
// last one-kernel step (marked at tstep[q] with clock - 1) raises the rebuild
// flag, the markers are cleared and the step clock slot takes the clock. With
// `out` (pinned host memory), the resulting status is also copied there.
__global__ void k_step_slots(DevStatus *st, int q, DevStatus *out) {
    const int64_t c = st->clock;
    if (st->tstep[q] == c - 1) st->rebuild = 1;
    st->tstep[0] = st->tstep[1] = INT64_MIN;
    st->cslot[q] = c;
    if (out) *out = *st;
}

// A copy of the status into pinned host memory (the snapshot the host reads
// after the event recorded behind it).
__global__ void k_snapshot(const DevStatus *st, DevStatus *out) { *out = *st; }

// Step budget of a simulation: freeze once `clock` reaches `target`.
__global__ void k_set_target(DevStatus *st, int64_t target) {
    st->target = target;
    st->stop = st->clock >= target ? 1 : 0;
}

// Step opening of an ensemble: the freeze decision of every replica (its
// rebuild flag read with acquire order, so that lists completed by a build on
// another stream are visible to the kernels that follow) and the clock tick.
__global__ void k_gate(DevStatus *st, int n, int flip) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int32_t rb = __hip_atomic_load(&st[r].rebuild, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const int32_t fr = rb | st[r].stop;
    st[r].update = fr;
    if (fr == 0) {
        st[r].clock += 1;
        if (flip) st[r].cur ^= 1;  // one-kernel steps: the step writes the other buffers
    }
}

// Buffers of an ensemble's one-kernel step: the device table names a
// replica's parity-0 buffers x, y, lab, rmax, rmin and its parity-1 ones xn,
// ...; k_gate has set `cur` to the parity the step writes, so the step reads
// the other one.
template <typename DP>
__device__ __forceinline__ DP step_buffers(const DP &t) {
    DP d = t;
    if (t.st->cur == 0) {
        d.x = t.xn;
        d.xn = t.x;
        d.y = t.yn;
        d.yn = t.y;
        d.lab = t.labn;
        d.labn = t.lab;
        d.rmax = t.rmaxn;
        d.rmaxn = t.rmax;
        d.rmin = t.rminn;
        d.rminn = t.rmin;
    }
    return d;
}

// After an ensemble run of one-kernel steps the host names each replica's
// current buffers as its first ones: every parity back to 0.
__global__ void k_clear_parity(DevStatus *st, int n) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) st[r].cur = 0;
}

// Start of a list build on the side stream: the neighbour kernel's maximum.
__global__ void k_begin_build(DevStatus *st) { st->overflow = 0; }

// End of a list build on the side stream: count it and release the replica
// (rebuild = 0, release order) unless the list overflowed its capacity or
// holds more neighbours than the forces kernel in use accepts (`limit`); the
// host then regrows or switches kernels.
__global__ void k_finish_build(DevStatus *st, int32_t cap, int32_t limit) {
    const int32_t n = st->overflow;
    st->builds += 1;
    const int32_t keep = (n > cap || n >= limit) ? 1 : 0;
    __hip_atomic_store(&st->rebuild, keep, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Step budgets of the n replicas of an ensemble run: target = clock + steps
// (steps < 0: none) and the matching `stop` flag.
__global__ void k_ensemble_targets(DevStatus *st, int n, int64_t steps) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t tg = steps < 0 ? INT64_MAX : st[r].clock + steps;
    st[r].target = tg;
    st[r].stop = st[r].clock >= tg ? 1 : 0;
}

// ------------------------------------------------------------ noise / misc

// noise_kernel! (src/random.jl:113-121): column i = the first three normals
// of philox_normals(T, key, i, draw).
template <typename T>
__global__ __launch_bounds__(256) void k_noise(int64_t n, uint32_t k0, uint32_t k1, uint64_t draw, T *out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    T z1, z2, z3, z4;
    philox_pair<T>(k0, k1, (uint32_t)(i + 1), draw, 0, z1, z2);
    philox_pair<T>(k0, k1, (uint32_t)(i + 1), draw, 1, z3, z4);
    out[3 * i] = z1;
    out[3 * i + 1] = z2;
    out[3 * i + 2] = z3;
}

__global__ __launch_bounds__(256) void k_philox(int64_t n, const uint32_t *ctr, const uint32_t *key,
                                                uint32_t *out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t c0 = ctr[4 * i], c1 = ctr[4 * i + 1], c2 = ctr[4 * i + 2], c3 = ctr[4 * i + 3];
    philox4x32(c0, c1, c2, c3, key[2 * i], key[2 * i + 1]);
    out[4 * i] = c0;
    out[4 * i + 1] = c1;
    out[4 * i + 2] = c2;
    out[4 * i + 3] = c3;
}

// bond_kernel! (src/energies/bonds.jl:98-110) for HarmonicBond terms
// (:25-29): gx, gy += the sum over the partners j of dE/dDelta at Delta = the
// minimum image of r_i - r_j, the terms summed as the right fold of
// _bond_gradient (:80-85). Runs after the forces kernel of a step; with
// use_status it skips a frozen step like the other kernels.
template <typename T, int UB, typename PP, typename DP>
__device__ __forceinline__ void k_bonds_impl(const PP &p, const DP &d, int use_status, int64_t blk) {
#pragma clang fp contract(off)
    const int64_t i = blk * (UB ? UB : (int)blockDim.x) + threadIdx.x;
    if (i >= p.N || d.bcount == nullptr) return;
    if (use_status && d.st->update != 0) return;
    const int n = d.bcount[i];
    if (n == 0) return;
    const int ti = d.type[i];
    const double xi = d.x[i], yi = d.y[i];
    T g0 = T(0), g1 = T(0);
    for (int q = 0; q < n; q++) {
        const int64_t j = d.bidx[(int64_t)q * p.N + i];
        const T dx = (T)min_image(xi - d.x[j], p.Lx);
        const T dy = (T)min_image(yi - d.y[j], p.Ly);
        const int tj = d.type[j];
        T a0 = T(0), a1 = T(0);
        for (int t = p.nbondt - 1; t >= 0; t--) {
            T f = T(0);
            if (((p.bdmask[t] >> ti) & 1u) && ((p.bdmask[t] >> tj) & 1u)) {
                const T r = sqrt(dx * dx + dy * dy);
                f = r > T(0) ? p.bdk[t] * (T(1) - p.bdr0[t] / r) : (p.bdr0[t] == T(0) ? p.bdk[t] : T(0));
            }
            a0 = f * dx + a0;
            a1 = f * dy + a1;
        }
        g0 += a0;
        g1 += a1;
    }
    d.gx[i] += g0;
    d.gy[i] += g1;
}

template <typename T>
__global__ __launch_bounds__(256) void k_bonds(const Params<T> p, const Ptrs<T> d, int use_status) {
    k_bonds_impl<T, 256>(p, d, use_status, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(256) void k_bonds_ens(const Params<T> *ps, const Ptrs<T> *ds, const Tile *tiles,
                                                   int use_status) {
    const Tile tl = load_uniform(tiles, blockIdx.x);
    if (tl.rep < 0) return;
    const FCC_CONST Params<T> &p = as_const(ps)[tl.rep];
    const FCC_CONST Ptrs<T> &d = as_const(ds)[tl.rep];
    k_bonds_impl<T, 0>(p, d, use_status, tl.blk);
}

// Perimeter and area sums of body i (src/observables.jl:6-14); given R, also
// the contours R, R', R'' at its nodes (nθ x N, column-major).
template <typename T, int KM>
__device__ __forceinline__ void contour_sums(const Params<T> &p, const Ptrs<T> &d, int64_t i, T *R, T *dR, T *ddR,
                                             T &per, T &area) {
    constexpr int NCM = 2 * KM + 1;
    const FCC_CONST T *S = as_const(d.S);
    T l[NCM];
    load_row<T, KM>(l, d.lab + i * p.ldc, p.K);
    const T r0 = d.R0[i];
    T ps = T(0), as = T(0);
    for (int k = 0; k < p.nth; k++) {
        T r, dr, ddr;
        synth<T, KM>(S, p.nth, k, l, p.K, true, r0, r, dr, ddr);
        if (R) {
            R[i * p.nth + k] = r;
            dR[i * p.nth + k] = dr;
            ddR[i * p.nth + k] = ddr;
        }
        ps += sqrt(r * r + dr * dr);
        as += r * r;
    }
    per = ps * p.dphi;
    area = as * (p.dphi / T(2));
}

template <typename T, int KM>
__global__ __launch_bounds__(256) void k_geometry(const Params<T> p, const Ptrs<T> d, T *R, T *dR, T *ddR,
                                                  T *per, T *area) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= p.N) return;
    T ps, as;
    contour_sums<T, KM>(p, d, i, R, dR, ddR, ps, as);
    if (per) {
        per[i] = ps;
        area[i] = as;
    }
}

}  // namespace fcc
