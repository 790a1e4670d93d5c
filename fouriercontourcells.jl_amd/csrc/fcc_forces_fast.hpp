// fcc_forces_fast.hpp — the forces kernels of the hot path (FP32 / FP64 VALU):
// k_forces_fast2 (two lanes per body, nθ = 40) and k_forces_fast (one lane per
// body, any nθ ≤ 64).
//
// Same work as k_forces (body terms, Overlap contact, pullback; src/gradients.jl,
// src/energies/contact.jl) with the shape of the hot configurations fixed at
// compile time:
//  - K (shape modes) is a template constant: no per-mode guards;
//  - Tension and PerimeterElasticity share the integrand √(R² + R'²)
//    (src/energies/body_terms.jl:12,33), so one perimeter integral and one
//    combined weight serve every such term; Bending has its own;
//  - the contact density is kept as per-node hit counts, one bit-sliced
//    counter (NP bit planes over the nθ nodes) per Overlap coupling class, in
//    registers: P_k = Σ_class J_class · count_class(k). A node's count is at
//    most the body's neighbour count, so the host uses this kernel only
//    while every list holds fewer than 2^NP entries;
//  - Overlap needs only the sign of d = R_j(ψ) - |u - r_j|: a node closer to
//    r_j than the lower bound rmin_j of R_j (minus a margin far above the
//    rounding of the evaluation) is inside without evaluating R_j(ψ); R_j(ψ)
//    is evaluated, by complex Horner (4 FMA per mode), only for nodes in the
//    shell rmin_j ≤ |u - r_j| ≤ rmax_j, so every decision equals the
//    reference's;
//  - only nodes inside the angular window that can reach a neighbour's
//    bounding disc are tested (window from polynomial asin and atan with a
//    safety margin far above their errors); the hit masks are node-indexed;
//  - each neighbour's position, bounds and lab row are loaded together, the
//    next neighbour's index one iteration ahead, and workgroups take their
//    tiles in XCD-grouped order;
//  - FP32 uses the hardware sqrt / rsqrt / rcp (≤ 1 ulp).
#pragma once

#include "fcc_kernels.hpp"

namespace fcc {

constexpr int NPLANE = 5;  // bit planes of the contact hit counters (counts < 32)

__device__ __forceinline__ float f_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double f_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float f_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double f_rsqrt(double x) { return 1.0 / sqrt(x); }
__device__ __forceinline__ float f_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double f_rcp(double x) { return 1.0 / x; }

// atan on [0, 1], Abramowitz & Stegun 4.4.49 (|error| ≤ 1e-5).
__device__ __forceinline__ float atan_unit(float t) {
    const float t2 = t * t;
    return t * (0.9998660f + t2 * (-0.3302995f + t2 * (0.1801410f + t2 * (-0.0851330f + t2 * 0.0208351f))));
}

// atan2 from atan_unit by octant reduction (|error| ≲ 2e-5 rad).
__device__ __forceinline__ float fast_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    float a = atan_unit(mn * __builtin_amdgcn_rcpf(mx));
    if (ay > ax) a = 1.57079633f - a;
    if (x < 0.f) a = 3.14159265f - a;
    return y < 0.f ? -a : a;
}

// asin on [0, 1], Abramowitz & Stegun 4.4.45 (|error| ≤ 5e-5; 6.8e-5 with
// FP32 rounding, checked on a 2·10⁶-point grid).
__device__ __forceinline__ float asin_unit(float x) {
    return 1.5707963f -
           __builtin_amdgcn_sqrtf(1.f - x) * (1.5707288f + x * (-0.2121144f + x * (0.0742610f + x * -0.0187293f)));
}

// Nodes [k0, k0 + wlen) (mod nθ) whose direction from body i lies within the
// half-angle asin(lim / D) of the direction to body j: no other node can be
// within lim of r_j (whatever its radius). Full circle if D ≤ lim. The 2e-4
// relative widening of lim² covers the reciprocal's rounding (1 ulp), the
// 2e-3 rad margin the polynomial asin and atan errors.
__device__ __forceinline__ void node_window(float dx, float dy, float lim2, float D2, float inv_dphi, int nth,
                                            int &k0, int &wlen) {
    k0 = 0;
    wlen = nth;
    const float sd2 = lim2 * 1.0002f * f_rcp(D2);
    if (!(sd2 < 1.f)) return;
    const float delta = asin_unit(__builtin_amdgcn_sqrtf(sd2)) + 2e-3f;
    const float psi = fast_atan2(-dy, -dx);
    const int klo = (int)floorf((psi - delta) * inv_dphi);
    const int khi = (int)ceilf((psi + delta) * inv_dphi);
    if (khi - klo + 1 >= nth) return;
    wlen = khi - klo + 1;
    // ψ ∈ [-π, π] and δ ≤ π/2 + 2e-3 give klo ∈ [-3nθ/4 - 1, nθ/2]: one
    // wrap brings it into [0, nθ)
    k0 = klo < 0 ? klo + nth : klo;
}

// Bit k of a 64-bit node mask, k uniform over the wave.
__device__ __forceinline__ uint32_t node_bit(uint64_t m, int k) {
    return k < 32 ? __builtin_amdgcn_ubfe((uint32_t)m, (uint32_t)k, 1u)
                  : __builtin_amdgcn_ubfe((uint32_t)(m >> 32), (uint32_t)(k - 32), 1u);
}

// Smallest squared distance for the direction of a node-to-centre offset:
// an offset of exactly zero takes z = 0, where the contour polynomial is a0 > 0,
// so the node counts as inside, as with the reference's ψ = 0 (R(0) > 0).
template <typename T>
__device__ __forceinline__ T tiny_d2() { return sizeof(T) == 4 ? T(1e-30) : T(1e-300); }

// R_j(ψ)/R0 = a0 + Σ_{m=1..K} αm cos (m+1)ψ + βm sin (m+1)ψ as Re Σ cn z^n,
// z = (c1, s1), by complex Horner (src/geometry.jl:147-156 evaluates the
// same sum by rotating (cos nψ, sin nψ)).
template <typename T, int K>
__device__ __forceinline__ T contour_horner(const T (&l)[2 * K + 1], T zr, T zi) {
    T sr = l[K], si = -l[2 * K];
#pragma unroll
    for (int n = K; n >= 2; n--) {
        const T nr = fma(sr, zr, fma(-si, zi, l[n - 1]));
        const T ni = fma(sr, zi, fma(si, zr, -l[K + n - 1]));
        sr = nr;
        si = ni;
    }
    const T nr = sr * zr - si * zi;
    const T ni = sr * zi + si * zr;
    return fma(nr, zr, fma(-ni, zi, l[0]));
}

// Contour R, R', R'' at node k from the compact basis rows [a0, cos.., sin..].
template <typename T, int K>
__device__ __forceinline__ void synth_k(const FCC_CONST T *S, int nth, int ldc, int k, const T (&l)[2 * K + 1],
                                        bool need_dd, T r0, T &R, T &dR, T &ddR) {
    const FCC_CONST T *C = S + (size_t)k * ldc;
    const FCC_CONST T *D = S + (size_t)(nth + k) * ldc;
    T r = l[0], dr = T(0);
#pragma unroll
    for (int c = 1; c <= 2 * K; c++) {
        r = fma(C[c], l[c], r);
        dr = fma(D[c], l[c], dr);
    }
    R = r0 * r;
    dR = r0 * dr;
    ddR = T(0);
    if (need_dd) {
        const FCC_CONST T *DD = S + (size_t)(2 * nth + k) * ldc;
        T ddr = T(0);
#pragma unroll
        for (int c = 1; c <= 2 * K; c++) ddr = fma(DD[c], l[c], ddr);
        ddR = r0 * ddr;
    }
}

template <typename T, int NTH_, int K, int NCLS, typename PP, typename DP>
__device__ __forceinline__ void k_forces_fast_impl(const PP &p, const DP &d, int use_status,
                                                    int debug, int64_t blk) {
    constexpr int NC = 2 * K + 1;
    constexpr int NTHM = NTH_ ? NTH_ : FUSED_MAX_THETA;
    __shared__ T sR[NTHM][64];
    __shared__ T scs[2][NTHM];
    const int nth = NTH_ ? NTH_ : p.nth;
    const int lane = threadIdx.x;
    for (int k = lane; k < nth; k += 64) {
        scs[0][k] = d.cs[2 * k];
        scs[1][k] = d.cs[2 * k + 1];
    }
    __syncthreads();
    const int64_t i = (int64_t)blk * 64 + lane;
    if (forces_frozen(use_status, d.st, blk, lane)) return;
    if (i >= p.N) return;
    const FCC_CONST T *S = as_const(d.Sc);
    const FCC_CONST T *CS = as_const(d.cs);
    const int ldc = p.ldc;
    constexpr int W = row_width<NC, T>();
    T lab[NC];
    {
        T a[W];
        row_load<T, W>(a, d.lab + i * ldc);
#pragma unroll
        for (int c = 0; c < NC; c++) lab[c] = a[c];
    }
    const T r0 = d.R0[i];
    const int ti = d.type[i];
    const uint32_t tbit = 1u << ti;
    const bool need_dd = p.need_dd != 0;
    const bool contact = p.ncontact != 0;

    // pass 1: perimeter integral (Tension, PerimeterElasticity) and Bending integral
    T Qp = T(0), Qb = T(0);
    for (int k = 0; k < nth; k++) {
        T R, dR, ddR;
        synth_k<T, K>(S, nth, ldc, k, lab, need_dd, r0, R, dR, ddR);
        if (contact) sR[k][lane] = R;
        Qp += f_sqrt(R * R + dR * dR);
        if (need_dd) Qb += term_q<T>(FCC_TERM_BENDING, R, dR, ddR);
    }
    T Wp = T(0), Wb = T(0), Et[MAXB];
#pragma unroll
    for (int b = 0; b < MAXB; b++) {
        Et[b] = T(0);
        if (b < p.nbody && (p.bmask[b] & tbit)) {
            T dF;
            const bool bend = p.bkind[b] == FCC_TERM_BENDING;
            Et[b] = term_E<T>(p.bkind[b], p.bp0[b], p.bp1[b], (bend ? Qb : Qp) * p.dphi, r0, dF);
            if (bend)
                Wb += dF * p.dphi;
            else
                Wp += dF * p.dphi;
        }
    }
    T E = T(0);
#pragma unroll
    for (int b = MAXB - 1; b >= 0; b--)
        if (b < p.nbody) E = Et[b] + E;

    // contact: hit counts per node and coupling class
    uint64_t pl[NCLS > 0 ? NCLS : 1][NPLANE];
#pragma unroll
    for (int c = 0; c < (NCLS > 0 ? NCLS : 1); c++)
#pragma unroll
        for (int q = 0; q < NPLANE; q++) pl[c][q] = 0;
    if (contact) {
        const double xi = d.x[i], yi = d.y[i];
        const T rmi = d.rmax[i];
        const int pi = d.parent[i];
        const T Lxf = (T)p.Lx, Lyf = (T)p.Ly;
        const bool wrap2 = p.small_box != 0;
        const float inv_dphi = (float)(1.0 / (double)p.dphi);
        int cnt = d.ncount[i];
        if (cnt > d.cap) cnt = d.cap;
        // The neighbour's position, bounds and lab row are loaded together
        // (one round trip per neighbour) and the next neighbour's index is
        // fetched while this one is processed.
        const int32_t *nrow = d.nidx + i;
        int jn = cnt > 0 ? nrow[0] : 0;
        for (int q = 0; q < cnt; q++) {
            const int j = jn;
            const double xj = d.x[j], yj = d.y[j];
            const T rmaxj = d.rmax[j], rminj = d.rmin[j], r0j = d.R0[j];
            T labj[W];
            row_load<T, W>(labj, d.lab + (int64_t)j * ldc);
            if (q + 1 < cnt) jn = nrow[(int64_t)(q + 1) * p.N];
            int cls = 0;
            if (!p.cls_uniform) {
                cls = p.ccls[pi == d.parent[j] ? 1 : 0][ti][d.type[j]];
                if (cls < 0) continue;  // zero coupling: no contribution
            }
            const T dxij = (T)min_image_nb(xi - xj, p.Lx, p.invLx, wrap2);
            const T dyij = (T)min_image_nb(yi - yj, p.Ly, p.invLy, wrap2);
            const T lim = rmaxj + p.range;
            const T lim2 = lim * lim;
            const T D2 = dxij * dxij + dyij * dyij;
            const T reach = (rmi + lim) * T(1.0001) + T(1e-6);
            if (D2 > reach * reach) continue;
            int k0, wlen;
            node_window((float)dxij, (float)dyij, (float)lim2, (float)D2, inv_dphi, nth, k0, wlen);
            // sure-inside radius: R_j(ψ) ≥ rmin_j exactly, and the evaluated
            // R_j differs from the exact one by far less than the margin
            T rin = rminj * (T(1) -(sizeof(T) == 8 ? T(1e-12) : T(2e-5))) - (sizeof(T) == 8 ? T(1e-13) : T(1e-6));
            const T rin2 = rin > T(0) ? rin * rin : T(-1);
            // node-indexed masks (bit k = node k): within lim of r_j and not
            // surely inside (mask); surely inside, then hits (hit)
            uint64_t mask = 0, hit = 0;
            int k = k0;
            for (int t = 0; t < wlen; t++) {
                const T Rk = sR[k][lane];
                T dx = dxij + Rk * scs[0][k];
                T dy = dyij + Rk * scs[1][k];
                if (wrap2) {
                    dx = min_image(dx, Lxf);
                    dy = min_image(dy, Lyf);
                }
                const T d2 = dx * dx + dy * dy;
                const uint64_t bit = 1ull << k;
                if (d2 <= lim2) mask |= bit;
                if (d2 < rin2) hit |= bit;
                if (++k == nth) k = 0;
            }
            mask &= ~hit;
            if (!mask) goto count_hits;
            {
            T lj[NC];
#pragma unroll
            for (int c = 0; c < NC; c++) lj[c] = labj[c];
            while (mask) {
                const int kk = __builtin_ctzll(mask);
                mask &= mask - 1;
                const T Rk = sR[kk][lane];
                T dx = dxij + Rk * scs[0][kk];
                T dy = dyij + Rk * scs[1][kk];
                if (wrap2) {
                    dx = min_image(dx, Lxf);
                    dy = min_image(dy, Lyf);
                }
                const T d2 = fmax(dx * dx + dy * dy, tiny_d2<T>());
                const T rs = f_rsqrt(d2);
                const T dist = d2 * rs;
                const T cp = dx * rs, sp = dy * rs;
                if (r0j * contour_horner<T, K>(lj, cp, sp) - dist >= T(0)) hit |= 1ull << kk;
            }
            }
        count_hits:
#pragma unroll
            for (int c = 0; c < (NCLS > 0 ? NCLS : 1); c++) {
                uint64_t m = cls == c ? hit : 0;
#pragma unroll
                for (int qq = 0; qq < NPLANE; qq++) {
                    const uint64_t carry = pl[c][qq] & m;
                    pl[c][qq] ^= m;
                    m = carry;
                }
            }
        }
    }

    // pass 2: contour gradients, contact force / torque, pullback
    T gl[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) gl[c] = T(0);
    T fx = T(0), fy = T(0), tt = T(0);
    constexpr int NCL = NCLS > 0 ? NCLS : 1;
    T Jd[NCL];
#pragma unroll
    for (int c = 0; c < NCL; c++) Jd[c] = p.Jcls[c] * p.dphi;
    // Planes 2.. hold counts of 4 or more (a node inside four neighbours):
    // they are read only when some lane of the wave has one.
    bool hi = false;
#pragma unroll
    for (int c = 0; c < NCL; c++)
#pragma unroll
        for (int q = 2; q < NPLANE; q++) hi |= pl[c][q] != 0;
    const bool hi_any = __ballot(hi) != 0;
    for (int k = 0; k < nth; k++) {
        T R, dR, ddR;
        synth_k<T, K>(S, nth, ldc, k, lab, need_dd, r0, R, dR, ddR);
        if (contact) R = sR[k][lane];
        const T rs = f_rsqrt(R * R + dR * dR);
        T GR = Wp * (R * rs), GD = Wp * (dR * rs), GDD = T(0);
        if (need_dd) {
            T j0, j1, j2;
            term_J<T>(FCC_TERM_BENDING, R, dR, ddR, j0, j1, j2);
            GR += Wb * j0;
            GD += Wb * j1;
            GDD = Wb * j2;
        }
        if (contact) {
            T pp = T(0), g = T(0);
#pragma unroll
            for (int c = 0; c < NCL; c++) {
                uint32_t n = node_bit(pl[c][0], k) | (node_bit(pl[c][1], k) << 1);
                if (hi_any) {
#pragma unroll
                    for (int qq = 2; qq < NPLANE; qq++) n |= node_bit(pl[c][qq], k) << qq;
                }
                pp += Jd[c] * (T)n;
                if (debug) g += p.Jcls[c] * (T)n;
            }
            const T c = CS[2 * k], s = CS[2 * k + 1];
            GR += pp * R;
            fx += pp * (R * c + dR * s);
            fy += pp * (R * s - dR * c);
            tt += (-pp) * R * dR;
            if (debug) d.P[i * nth + k] = g;
        }
        const FCC_CONST T *C = S + (size_t)k * ldc;
        const FCC_CONST T *D = S + (size_t)(nth + k) * ldc;
        gl[0] += GR;
#pragma unroll
        for (int c = 1; c < NC; c++) gl[c] = fma(C[c], GR, fma(D[c], GD, gl[c]));
        if (need_dd) {
            const FCC_CONST T *DD = S + (size_t)(2 * nth + k) * ldc;
#pragma unroll
            for (int c = 1; c < NC; c++) gl[c] = fma(DD[c], GDD, gl[c]);
        }
        if (debug) {
            d.GR[i * nth + k] = GR;
            d.GD[i * nth + k] = GD;
            d.GDD[i * nth + k] = GDD;
        }
    }
    {
        T a[W];
#pragma unroll
        for (int c = 0; c < W; c++) a[c] = c < NC ? gl[c] : T(0);
        row_store<T, W>(a, d.glab + i * ldc);
    }
    d.gx[i] = fx;
    d.gy[i] = fy;
    d.gth[i] = tt;
    d.Ebody[i] = E;
    if (d.th0) d.th0[i] = d.th[i];  // parent drag: θ at the start of the step
}

template <typename T, int NTH_, int K, int NCLS>
__global__ __launch_bounds__(64) void k_forces_fast(const Params<T> p, const Ptrs<T> d, int use_status,
                                                    int debug) {
    k_forces_fast_impl<T, NTH_, K, NCLS>(p, d, use_status, debug, xcd_tile(blockIdx.x, gridDim.x));
}

template <typename T, int NTH_, int K, int NCLS>
__global__ __launch_bounds__(64) void k_forces_fast_ens(const Params<T> *ps, const Ptrs<T> *ds, const Tile *tiles,
                                                        int use_status, int debug) {
    const Tile tl = load_uniform(tiles, blockIdx.x);
    if (tl.rep < 0) return;
    const FCC_CONST Params<T> &p = as_const(ps)[tl.rep];
    const FCC_CONST Ptrs<T> &d = as_const(ds)[tl.rep];
    k_forces_fast_impl<T, NTH_, K, NCLS>(p, d, use_status, debug, tl.blk);
}

// Bit-sliced sum of two per-node counters (NPLANE planes each).
__device__ __forceinline__ void planes_add(uint64_t (&a)[NPLANE], const uint64_t (&b)[NPLANE]) {
    uint64_t c = 0;
#pragma unroll
    for (int q = 0; q < NPLANE; q++) {
        const uint64_t s = a[q] ^ b[q] ^ c;
        c = (a[q] & b[q]) | (c & (a[q] ^ b[q]));
        a[q] = s;
    }
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    const uint32_t lo = __shfl_xor((uint32_t)v, m), hi = __shfl_xor((uint32_t)(v >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}

// Mode n's (cos, sin) pair of a coefficient row, in place, for the turn by
// nα with nα a multiple r of π/2 (r = 0..3): cos nα, sin nα ∈ {0, ±1}, so the
// rotation only swaps and negates (exact). `inv` applies the inverse turn.
template <typename T>
__device__ __forceinline__ void quarter_turn(T &c, T &s, int r, bool inv) {
    const T a = c, b = s;
    if (inv && (r & 1)) r ^= 2;
    c = r == 0 ? a : r == 1 ? b : r == 2 ? -a : -b;
    s = r == 0 ? b : r == 1 ? -a : r == 2 ? -b : a;
}

// The same work with LPB = 2 or 4 lanes per body (nθ = 40): 64 / LPB bodies
// per wave, lane (b, h) takes nodes 40h/LPB .. 40(h+1)/LPB - 1 of body b in
// the node passes and neighbour slots q ≡ h (mod LPB) over the whole circle in
// the contact pass (the mixed-precision kernel's layout for LPB = 2; LPB = 4
// doubles the waves of small systems). Node j + 40h/LPB sits at φ_j + hα,
// α = 2π/LPB, where the basis rows are those of node j with each mode's (cos,
// sin) pair turned by n·hα, a multiple of π/2: lane group h turns its lab row
// once, so all groups read the same wave-uniform basis rows through the scalar
// cache, and its pullback sums take the inverse turn before the groups are
// added.
//
// NW waves per workgroup (1, or NW = LPB: 64 bodies), wave w on the bodies of
// tile wg·NW + w; with NW = LPB the workgroup is a one-kernel step (the
// pullback rows and centre force / torque go to the waves' LDS slices, then
// wave 0 updates the 64 bodies, fused_update).
template <typename T, int K, int NCLS, int LPB, int NW, typename PP, typename DP>
__device__ __forceinline__ void k_forces_fast2_impl(const PP &p, const DP &d, int use_status, int debug,
                                                     int64_t wg, int q = 0) {
    static_assert(LPB == 2 || LPB == 4, "two or four lanes per body");
    static_assert(NW == 1 || NW == LPB, "one wave, or 64 bodies per workgroup");
    constexpr int NTH = 40, NB = 64 / LPB, NH2 = NTH / LPB;
    constexpr int NC = 2 * K + 1;
    constexpr int NCL = NCLS > 0 ? NCLS : 1;
    constexpr int W = row_width<NC, T>();
    static_assert(NB * W + 3 * NB <= NTH * NB, "pullback rows fit the wave's LDS slice");
    __shared__ T sRw[NW][NTH][NB];
    __shared__ T scs[2][NTH];
    const int lane = threadIdx.x & 63, w = NW == 1 ? 0 : (int)(threadIdx.x >> 6);
    T(*sR)[NB] = sRw[w];
    const int b = lane & (NB - 1), h = lane / NB;
    for (int k = threadIdx.x; k < NTH; k += 64 * NW) {
        scs[0][k] = d.cs[2 * k];
        scs[1][k] = d.cs[2 * k + 1];
    }
    const int64_t blk = wg * NW + w;
    const int64_t i = (int64_t)blk * NB + b;
    int64_t clk = 0;  // one-kernel step: its clock
    if constexpr (NW > 1) {
        if (use_status == 2) {  // ensemble step: k_gate decided and ticked the clock
            if (d.st->update != 0) return;
            clk = d.st->clock - 1;
        } else if (step_open(d.st, q, use_status == 4, blk, lane, clk)) {
            return;  // uniform over the grid
        }
    } else {
        if (forces_frozen(use_status, d.st, blk, lane)) return;  // uniform over the grid
    }
    const bool live = i < p.N;
    const int64_t ii = live ? i : 0;
    const FCC_CONST T *S = as_const(d.Sc);
    const int ldc = p.ldc;
    const int qt = h * (4 / LPB);  // quarter turns of this lane group's first node
    T lab[NC];
    {
        T a[W];
        row_load<T, W>(a, d.lab + ii * ldc);
#pragma unroll
        for (int c = 0; c < NC; c++) lab[c] = a[c];
#pragma unroll
        for (int m = 1; m <= K; m++) quarter_turn<T>(lab[m], lab[K + m], ((m + 1) * qt) & 3, false);
    }
    const T r0 = live ? d.R0[ii] : T(1);
    const int ti = live ? d.type[ii] : 0;
    const uint32_t tbit = 1u << ti;
    const bool need_dd = p.need_dd != 0;
    const bool contact = p.ncontact != 0;

    // pass 1: perimeter and bending integrals over both halves
    T Qp = T(0), Qb = T(0);
    for (int j = 0; j < NH2; j++) {
        T R, dR, ddR;
        synth_k<T, K>(S, NTH, ldc, j, lab, need_dd, r0, R, dR, ddR);
        if (contact) sR[h * NH2 + j][b] = R;
        Qp += f_sqrt(R * R + dR * dR);
        if (need_dd) Qb += term_q<T>(FCC_TERM_BENDING, R, dR, ddR);
    }
#pragma unroll
    for (int o = NB; o < 64; o <<= 1) {
        Qp += __shfl_xor(Qp, o);
        Qb += __shfl_xor(Qb, o);
    }
    T Wp = T(0), Wb = T(0), Et[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        Et[q] = T(0);
        if (q < p.nbody && (p.bmask[q] & tbit)) {
            T dF;
            const bool bend = p.bkind[q] == FCC_TERM_BENDING;
            Et[q] = term_E<T>(p.bkind[q], p.bp0[q], p.bp1[q], (bend ? Qb : Qp) * p.dphi, r0, dF);
            if (bend)
                Wb += dF * p.dphi;
            else
                Wp += dF * p.dphi;
        }
    }
    T E = T(0);
#pragma unroll
    for (int q = MAXB - 1; q >= 0; q--)
        if (q < p.nbody) E = Et[q] + E;
    __syncthreads();  // sR, scs complete

    // contact: hit counts per node and coupling class
    uint64_t pl[NCL][NPLANE];
#pragma unroll
    for (int c = 0; c < NCL; c++)
#pragma unroll
        for (int q = 0; q < NPLANE; q++) pl[c][q] = 0;
    if (contact && live) {
        const double xi = d.x[ii], yi = d.y[ii];
        const T rmi = d.rmax[ii];
        const int pi = d.parent[ii];
        const T Lxf = (T)p.Lx, Lyf = (T)p.Ly;
        const bool wrap2 = p.small_box != 0;
        const float inv_dphi = (float)(1.0 / (double)p.dphi);
        int cnt = d.ncount[ii];
        if (cnt > d.cap) cnt = d.cap;
        // the neighbour's position, bounds and lab row in one round trip, the
        // next neighbour's index one iteration ahead
        const int32_t *nrow = d.nidx + ii;
        int jn = h < cnt ? nrow[(int64_t)h * p.N] : 0;
        for (int q = h; q < cnt; q += LPB) {
            const int j = jn;
            const double xj = d.x[j], yj = d.y[j];
            const T rmaxj = d.rmax[j], rminj = d.rmin[j], r0j = d.R0[j];
            T labj[W];
            row_load<T, W>(labj, d.lab + (int64_t)j * ldc);
            if (q + LPB < cnt) jn = nrow[(int64_t)(q + LPB) * p.N];
            int cls = 0;
            if (!p.cls_uniform) {
                cls = p.ccls[pi == d.parent[j] ? 1 : 0][ti][d.type[j]];
                if (cls < 0) continue;  // zero coupling: no contribution
            }
            const T dxij = (T)min_image_nb(xi - xj, p.Lx, p.invLx, wrap2);
            const T dyij = (T)min_image_nb(yi - yj, p.Ly, p.invLy, wrap2);
            const T lim = rmaxj + p.range;
            const T lim2 = lim * lim;
            const T D2 = dxij * dxij + dyij * dyij;
            const T reach = (rmi + lim) * T(1.0001) + T(1e-6);
            if (D2 > reach * reach) continue;
            int k0, wlen;
            node_window((float)dxij, (float)dyij, (float)lim2, (float)D2, inv_dphi, NTH, k0, wlen);
            // sure-inside radius: R_j(ψ) ≥ rmin_j exactly, and the evaluated
            // R_j differs from the exact one by far less than the margin
            T rin = rminj * (T(1) - (sizeof(T) == 8 ? T(1e-12) : T(2e-5))) - (sizeof(T) == 8 ? T(1e-13) : T(1e-6));
            const T rin2 = rin > T(0) ? rin * rin : T(-1);
            // node-indexed masks (bit k = node k): within lim of r_j and not
            // surely inside (mask); surely inside, then hits (hit)
            uint64_t mask = 0, hit = 0;
            int k = k0;
            for (int t = 0; t < wlen; t++) {
                const T Rk = sR[k][b];
                T dx = dxij + Rk * scs[0][k];
                T dy = dyij + Rk * scs[1][k];
                if (wrap2) {
                    dx = min_image(dx, Lxf);
                    dy = min_image(dy, Lyf);
                }
                const T d2 = dx * dx + dy * dy;
                const uint64_t bit = 1ull << k;
                if (d2 <= lim2) mask |= bit;
                if (d2 < rin2) hit |= bit;
                if (++k == NTH) k = 0;
            }
            mask &= ~hit;
            if (mask) {
                T lj[NC];
#pragma unroll
                for (int c = 0; c < NC; c++) lj[c] = labj[c];
                while (mask) {
                    const int kk = __builtin_ctzll(mask);
                    mask &= mask - 1;
                    const T Rk = sR[kk][b];
                    T dx = dxij + Rk * scs[0][kk];
                    T dy = dyij + Rk * scs[1][kk];
                    if (wrap2) {
                        dx = min_image(dx, Lxf);
                        dy = min_image(dy, Lyf);
                    }
                    const T d2 = fmax(dx * dx + dy * dy, tiny_d2<T>());
                    const T rs = f_rsqrt(d2);
                    const T dist = d2 * rs;
                    const T cp = dx * rs, sp = dy * rs;
                    if (r0j * contour_horner<T, K>(lj, cp, sp) - dist >= T(0)) hit |= 1ull << kk;
                }
            }
#pragma unroll
            for (int c = 0; c < NCL; c++) {
                uint64_t m = cls == c ? hit : 0;
#pragma unroll
                for (int qq = 0; qq < 2; qq++) {
                    const uint64_t carry = pl[c][qq] & m;
                    pl[c][qq] ^= m;
                    m = carry;
                }
                // a carry into plane 2 (a node reaching a count of 4) is rare
                if (__ballot(m != 0)) {
#pragma unroll
                    for (int qq = 2; qq < NPLANE; qq++) {
                        const uint64_t carry = pl[c][qq] & m;
                        pl[c][qq] ^= m;
                        m = carry;
                    }
                }
            }
        }
    }
    // the lanes of a body: add the other groups' counters
#pragma unroll
    for (int o = NB; o < 64; o <<= 1)
#pragma unroll
        for (int c = 0; c < NCL; c++) {
            uint64_t other[NPLANE];
#pragma unroll
            for (int q = 0; q < NPLANE; q++) other[q] = shfl_xor64(pl[c][q], o);
            planes_add(pl[c], other);
        }
    // the lane's nodes of each counter plane at bits 0..NH2-1
    uint32_t plh[NCL][NPLANE];
#pragma unroll
    for (int c = 0; c < NCL; c++)
#pragma unroll
        for (int q = 0; q < NPLANE; q++) plh[c][q] = (uint32_t)(pl[c][q] >> (h * NH2));

    // pass 2: contour gradients, contact force / torque, pullback
    T gl[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) gl[c] = T(0);
    T fx = T(0), fy = T(0), tt = T(0);
    T Jd[NCL];
#pragma unroll
    for (int c = 0; c < NCL; c++) Jd[c] = p.Jcls[c] * p.dphi;
    // Planes 2.. hold counts of 4 or more (a node inside four neighbours):
    // they are read only when some lane of the wave has one.
    bool hi = false;
#pragma unroll
    for (int c = 0; c < NCL; c++)
#pragma unroll
        for (int q = 2; q < NPLANE; q++) hi |= plh[c][q] != 0;
    const bool hi_any = __ballot(hi) != 0;
    for (int j = 0; j < NH2; j++) {
        const int k = h * NH2 + j;
        T R, dR, ddR;
        synth_k<T, K>(S, NTH, ldc, j, lab, need_dd, r0, R, dR, ddR);
        if (contact) R = sR[k][b];
        const T rs = f_rsqrt(R * R + dR * dR);
        T GR = Wp * (R * rs), GD = Wp * (dR * rs), GDD = T(0);
        if (need_dd) {
            T j0, j1, j2;
            term_J<T>(FCC_TERM_BENDING, R, dR, ddR, j0, j1, j2);
            GR += Wb * j0;
            GD += Wb * j1;
            GDD = Wb * j2;
        }
        if (contact) {
            T pp = T(0), g = T(0);
#pragma unroll
            for (int c = 0; c < NCL; c++) {
                uint32_t n = __builtin_amdgcn_ubfe(plh[c][0], (uint32_t)j, 1u) |
                             (__builtin_amdgcn_ubfe(plh[c][1], (uint32_t)j, 1u) << 1);
                if (hi_any) {
#pragma unroll
                    for (int qq = 2; qq < NPLANE; qq++) n |= __builtin_amdgcn_ubfe(plh[c][qq], (uint32_t)j, 1u) << qq;
                }
                pp += Jd[c] * (T)n;
                if (debug) g += p.Jcls[c] * (T)n;
            }
            const T c = scs[0][k], s = scs[1][k];
            GR += pp * R;
            fx += pp * (R * c + dR * s);
            fy += pp * (R * s - dR * c);
            tt += (-pp) * R * dR;
            if (debug && live) d.P[ii * NTH + k] = g;
        }
        const FCC_CONST T *C = S + (size_t)j * ldc;
        const FCC_CONST T *D = S + (size_t)(NTH + j) * ldc;
        gl[0] += GR;
#pragma unroll
        for (int c = 1; c < NC; c++) gl[c] = fma(C[c], GR, fma(D[c], GD, gl[c]));
        if (need_dd) {
            const FCC_CONST T *DD = S + (size_t)(2 * NTH + j) * ldc;
#pragma unroll
            for (int c = 1; c < NC; c++) gl[c] = fma(DD[c], GDD, gl[c]);
        }
        if (debug && live) {
            d.GR[ii * NTH + k] = GR;
            d.GD[ii * NTH + k] = GD;
            d.GDD[ii * NTH + k] = GDD;
        }
    }
    // the group summed with the rows of nodes 0..NH2-1: inverse turn per mode
#pragma unroll
    for (int m = 1; m <= K; m++) quarter_turn<T>(gl[m], gl[K + m], ((m + 1) * qt) & 3, true);
#pragma unroll
    for (int o = NB; o < 64; o <<= 1) {
#pragma unroll
        for (int c = 0; c < NC; c++) gl[c] += __shfl_xor(gl[c], o);
        fx += __shfl_xor(fx, o);
        fy += __shfl_xor(fy, o);
        tt += __shfl_xor(tt, o);
    }
    if (live && h == 0) {
        T a[W];
#pragma unroll
        for (int c = 0; c < W; c++) a[c] = c < NC ? gl[c] : T(0);
        row_store<T, W>(a, d.glab + ii * ldc);
        d.gx[ii] = fx;
        d.gy[ii] = fy;
        d.gth[ii] = tt;
        d.Ebody[ii] = E;
        if (d.th0) d.th0[ii] = d.th[ii];  // parent drag: θ at the start of the step
    }
    if constexpr (NW > 1) {
        // the same row, force and torque into the wave's LDS slice (pass 2
        // has read its contour slots): body b at rows[b·W], the force and
        // torque after the NB rows
        T *rows = &sR[0][0];
        if (h == 0) {
#pragma unroll
            for (int c = 0; c < NC; c++) rows[b * W + c] = gl[c];
            rows[NB * W + 3 * b] = fx;
            rows[NB * W + 3 * b + 1] = fy;
            rows[NB * W + 3 * b + 2] = tt;
        }
        __syncthreads();
        if (w != 0) return;
        fused_update<T, K, false>(p, d, &sRw[0][0][0], NTH * NB, NB, wg * 64, lane, q, clk, use_status == 2);
    }
}

template <typename T, int K, int NCLS, int LPB>
__global__ __launch_bounds__(64) void k_forces_fast2(const Params<T> p, const Ptrs<T> d, int use_status, int debug) {
    k_forces_fast2_impl<T, K, NCLS, LPB, 1>(p, d, use_status, debug, xcd_tile(blockIdx.x, gridDim.x));
}

// One-kernel step of the plain path (two lanes per body, two waves and 64
// bodies per workgroup) reading the buffers of parity q; use_status 3, or 4
// for timing series.
template <typename T, int K, int NCLS>
__global__ __launch_bounds__(128) void k_step_fast2(const Params<T> p, const Ptrs<T> d, int use_status, int q) {
    k_forces_fast2_impl<T, K, NCLS, 2, 2>(p, d, use_status, 0, xcd_tile(blockIdx.x, gridDim.x), q);
}

// Ensemble form of the one-kernel step: tiles of 64 bodies, each replica's
// buffers by the parity k_gate left in its status (step_buffers).
template <typename T, int K, int NCLS>
__global__ __launch_bounds__(128) void k_step_fast2_ens(const Params<T> *ps, const Ptrs<T> *ds, const Tile *tiles) {
    const Tile tl = load_uniform(tiles, blockIdx.x);
    if (tl.rep < 0) return;
    const FCC_CONST Params<T> &p = as_const(ps)[tl.rep];
    const Ptrs<T> d = step_buffers(load_uniform(ds, tl.rep));
    k_forces_fast2_impl<T, K, NCLS, 2, 2>(p, d, 2, 0, tl.blk);
}

// The one-kernel step with four lanes per body (systems below
// FOUR_LANES_BELOW bodies): four waves of 16 bodies per workgroup.
template <typename T, int K, int NCLS>
__global__ __launch_bounds__(256) void k_step_fast4(const Params<T> p, const Ptrs<T> d, int use_status, int q) {
    k_forces_fast2_impl<T, K, NCLS, 4, 4>(p, d, use_status, 0, xcd_tile(blockIdx.x, gridDim.x), q);
}

template <typename T, int K, int NCLS>
__global__ __launch_bounds__(256) void k_step_fast4_ens(const Params<T> *ps, const Ptrs<T> *ds, const Tile *tiles) {
    const Tile tl = load_uniform(tiles, blockIdx.x);
    if (tl.rep < 0) return;
    const FCC_CONST Params<T> &p = as_const(ps)[tl.rep];
    const Ptrs<T> d = step_buffers(load_uniform(ds, tl.rep));
    k_forces_fast2_impl<T, K, NCLS, 4, 4>(p, d, 2, 0, tl.blk);
}

template <typename T, int K, int NCLS, int LPB>
__global__ __launch_bounds__(64) void k_forces_fast2_ens(const Params<T> *ps, const Ptrs<T> *ds, const Tile *tiles,
                                                         int use_status, int debug) {
    const Tile tl = load_uniform(tiles, blockIdx.x);
    if (tl.rep < 0) return;
    const FCC_CONST Params<T> &p = as_const(ps)[tl.rep];
    const FCC_CONST Ptrs<T> &d = as_const(ds)[tl.rep];
    k_forces_fast2_impl<T, K, NCLS, LPB, 1>(p, d, use_status, debug, tl.blk);
}

}  // namespace fcc
