// fcc_forces_mfma4.hpp — the mixed-precision forces kernel with four lanes per
// body, for systems below FOUR_LANES_BELOW bodies (where 32 bodies per wave
// leave two waves per SIMD and the kernel waits on its gathers).
//
// Layout: 16 bodies per wave. Lane l works on body b = l % 16 and quarter
// g = l / 16 of its contour (nodes 10g..10g+9 for nθ = 40). The synthesis and
// the pullback run on v_mfma_f32_16x16x16_{f16,bf16} tiles: the output of a
// tile holds, in lane l, rows 4g..4g+3 of column l % 16; with the basis rows
// permuted on the host, the lane's slots s = 4t + r (tile t, row r) are R
// (s < 10), R' (s < 20) and R'' (s < 30, with Bending) at its nodes. The same
// registers are the A operand of the pullback tiles (lane l supplies rows
// l % 16, K-elements 4g..4g+3), so the contour gradients that replace R, R',
// R'' feed the pullback without a shuffle. Contact as in the four-lane FP32
// kernel: lane group g takes neighbour slots q ≡ g (mod 4) over the whole
// circle, and the bit-sliced counters of the four groups are added with two
// butterfly steps. hi+lo splits and scales as in k_forces_mfma.
#pragma once

#include "fcc_forces_mfma.hpp"

namespace fcc {

typedef _Float16 h4_t __attribute__((ext_vector_type(4)));
typedef short s4_t __attribute__((ext_vector_type(4)));
typedef float f4v_t __attribute__((ext_vector_type(4)));

constexpr int NQ4 = NTH_M / 4;  // nodes per lane group

template <bool BF>
struct Mfma16;
template <>
struct Mfma16<false> {
    typedef h4_t v_t;
    static __device__ __forceinline__ f4v_t mfma(v_t a, v_t b, f4v_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ void split(const float (&x)[4], v_t &hi, v_t &lo) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            float v = x[s];
            asm volatile("" : "+v"(v));  // as in split8: the FP32 value, then its split
            const _Float16 h = (_Float16)v;
            hi[s] = h;
            lo[s] = (_Float16)(v - (float)h);
        }
    }
};
template <>
struct Mfma16<true> {
    typedef s4_t v_t;  // bf16 bit patterns
    static __device__ __forceinline__ f4v_t mfma(v_t a, v_t b, f4v_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ void split(const float (&x)[4], v_t &hi, v_t &lo) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            float v = x[s];
            asm volatile("" : "+v"(v));  // as in split8: the FP32 value, then its split
            const __bf16 h = (__bf16)v;
            const __bf16 l = (__bf16)(v - (float)h);
            hi[s] = __builtin_bit_cast(short, h);
            lo[s] = __builtin_bit_cast(short, l);
        }
    }
};

template <bool BF>
__device__ __forceinline__ typename Mfma16<BF>::v_t load_op4(const uint16_t *p) {
    typename Mfma16<BF>::v_t v;
    const uint2 q = *reinterpret_cast<const uint2 *>(p);
    __builtin_memcpy(&v, &q, 8);
    return v;
}

// Tiles of the four-lane layout: 5 (R, R' at 10 nodes) or 8 (R, R', R'').
template <int T>
constexpr int tiles4() { return T == 4 ? 8 : 5; }

// Synthesis: lane (b, g) supplies lab coefficients 4g..4g+3 of body b
// (pre-scaled by SYN_SCALE); acc[t] receives tile t of S·(SYN_SCALE·lab).
template <int K, int TT, bool BF>
__device__ __forceinline__ void synth4(const float *lab_row, bool live, int g, int lane, const uint16_t *asyn,
                                       f4v_t (&acc)[TT]) {
    typedef Mfma16<BF> MT;
    typedef typename MT::v_t v_t;
    constexpr int NC = 2 * K + 1;
    const float4 u0 = *reinterpret_cast<const float4 *>(lab_row + 4 * g);
    float x[4] = {u0.x, u0.y, u0.z, u0.w};
#pragma unroll
    for (int s = 0; s < 4; s++) x[s] = (live && 4 * g + s < NC) ? x[s] * SYN_SCALE : 0.f;
    v_t bh, bl;
    MT::split(x, bh, bl);
#pragma unroll
    for (int t = 0; t < TT; t++) {
        const v_t ah = load_op4<BF>(asyn + ((t * 2 + 0) * 64 + lane) * 4);
        const v_t al = load_op4<BF>(asyn + ((t * 2 + 1) * 64 + lane) * 4);
        f4v_t c = {};
        c = MT::mfma(ah, bh, c);
        c = MT::mfma(ah, bl, c);
        c = MT::mfma(al, bh, c);
        acc[t] = c;
    }
}

// Operand tables (built on the host, 16-bit values):
//  asyn[(t * 2 + part) * 64 + lane][4]   synthesis A (S rows), part 0 hi / 1 lo
//  bpull[(t * 2 + part) * 64 + lane][4]  pullback B (S rows of tile t), part 0 hi / 1 lo
//
// NW waves per workgroup (1 or 4), wave w on the 16 bodies of tile wg·NW + w.
// With NW = 4 the workgroup is a one-kernel step of 64 bodies: the waves
// leave their bodies' pullback rows and centre force / torque in their LDS
// slices, and wave 0 updates the 64 bodies (fused_update).
template <int K, int T, int NCLS, bool BF, int NW, typename PP, typename DP>
__device__ __forceinline__ void k_forces_mfma4_impl(const PP &p, const DP &d, const uint16_t *asyn,
                                                     const uint16_t *bpull, int use_status, int debug,
                                                     int64_t wg, int q = 0) {
    typedef Mfma16<BF> MT;
    typedef typename MT::v_t v_t;
    constexpr int NC = 2 * K + 1;
    constexpr int W = row_width<NC, float>();
    constexpr int TT = tiles4<T>();
    static_assert(NW == 1 || NW == 4, "one wave, or 64 bodies per workgroup");
    static_assert(16 * W + 3 * 16 <= NTH_M * 16, "pullback rows fit the wave's LDS slice");
    __shared__ float sRw[NW][NTH_M][16];  // R of each wave's 16 bodies for the contact pass
    __shared__ float scs[2][NTH_M];
    const int lane = threadIdx.x & 63, w = NW == 1 ? 0 : (int)(threadIdx.x >> 6);
    float(*sR)[16] = sRw[w];
    const int b = lane & 15, g = lane >> 4;
    for (int k = threadIdx.x; k < NTH_M; k += 64 * NW) {
        scs[0][k] = d.cs[2 * k];
        scs[1][k] = d.cs[2 * k + 1];
    }
    const int64_t blk = wg * NW + w;
    const int64_t i = (int64_t)blk * 16 + b;
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
    const float r0 = live ? d.R0[ii] : 1.f;
    const float u = r0 * (1.f / SYN_SCALE);  // R = u·slot
    const float inv_u = SYN_SCALE / r0;
    const float *lab_row = d.lab + ii * LDC_M;
    const int ti = live ? d.type[ii] : 0;
    const uint32_t tbit = 1u << ti;
    const bool contact = p.ncontact != 0;

    // ---- synthesis on MFMA
    f4v_t acc[TT];
    synth4<K, TT, BF>(lab_row, live, g, lane, asyn, acc);
    // local slot s = 4t + r → acc[t][r]; R/u at s = j, R'/u at 10 + j, R''/u at 20 + j
#define SLOT4(s) acc[(s) >> 2][(s)&3]

    // ---- pass 1: perimeter and bending integrals over the four groups
    float Qp = 0.f, Qb = 0.f;
#pragma unroll
    for (int j = 0; j < NQ4; j++) {
        const float a = SLOT4(j), ad = SLOT4(NQ4 + j);
        Qp += f_sqrt(a * a + ad * ad);
        if constexpr (T == 4) Qb += term_q<float>(FCC_TERM_BENDING, a, ad, SLOT4(2 * NQ4 + j));
        sR[g * NQ4 + j][b] = a * u;
    }
    Qp = Qp + __shfl_xor(Qp, 16);
    Qp = (Qp + __shfl_xor(Qp, 32)) * u;
    if constexpr (T == 4) {
        Qb = Qb + __shfl_xor(Qb, 16);
        Qb = (Qb + __shfl_xor(Qb, 32)) * inv_u;
    }
    float Wp = 0.f, Wb = 0.f, Et[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; q++) {
        Et[q] = 0.f;
        if (q < p.nbody && (p.bmask[q] & tbit)) {
            float dF;
            const bool bend = p.bkind[q] == FCC_TERM_BENDING;
            Et[q] = term_E<float>(p.bkind[q], p.bp0[q], p.bp1[q], (bend ? Qb : Qp) * p.dphi, r0, dF);
            if (bend)
                Wb += dF * p.dphi;
            else
                Wp += dF * p.dphi;
        }
    }
    float E = 0.f;
#pragma unroll
    for (int q = MAXB - 1; q >= 0; q--)
        if (q < p.nbody) E = Et[q] + E;
    __syncthreads();  // sR, scs complete

    // ---- contact: lane group g takes neighbour slots q ≡ g (mod 4), full circle
    uint64_t pl[NCLS][NPLANE];
#pragma unroll
    for (int c = 0; c < NCLS; c++)
#pragma unroll
        for (int q = 0; q < NPLANE; q++) pl[c][q] = 0;
    if (contact && live) {
        const double xi = d.x[ii], yi = d.y[ii];
        const float rmi = d.rmax[ii];
        const int pi = d.parent[ii];
        const float Lxf = (float)p.Lx, Lyf = (float)p.Ly;
        const bool wrap2 = p.small_box != 0;
        const float inv_dphi = (float)(1.0 / (double)p.dphi);
        int cnt = d.ncount[ii];
        if (cnt > d.cap) cnt = d.cap;
        // The neighbour's position, bounds and lab row are loaded together
        // (one round trip per neighbour) and the next neighbour's index is
        // fetched while this one is processed.
        const int32_t *nrow = d.nidx + ii;
        int jn = g < cnt ? nrow[(int64_t)g * p.N] : 0;
        for (int q = g; q < cnt; q += 4) {
            const int j = jn;
            const double xj = d.x[j], yj = d.y[j];
            const float rmaxj = d.rmax[j], rminj = d.rmin[j], r0j = d.R0[j];
            float labj[W];
            row_load<float, W>(labj, d.lab + (int64_t)j * LDC_M);
            if (q + 4 < cnt) jn = nrow[(int64_t)(q + 4) * p.N];
            int cls = 0;
            if (!p.cls_uniform) {
                cls = p.ccls[pi == d.parent[j] ? 1 : 0][ti][d.type[j]];
                if (cls < 0) continue;
            }
            const float dxij = (float)min_image_nb(xi - xj, p.Lx, p.invLx, wrap2);
            const float dyij = (float)min_image_nb(yi - yj, p.Ly, p.invLy, wrap2);
            const float lim = rmaxj + p.range;
            const float lim2 = lim * lim;
            const float D2 = dxij * dxij + dyij * dyij;
            const float reach = (rmi + lim) * 1.0001f + 1e-6f;
            if (D2 > reach * reach) continue;
            int k0, wlen;
            node_window(dxij, dyij, lim2, D2, inv_dphi, NTH_M, k0, wlen);
            float rin = rminj * (1.f - 2e-5f) - 1e-6f;
            const float rin2 = rin > 0.f ? rin * rin : -1.f;
            // node-indexed masks (bit k = node k): within lim of r_j and not
            // surely inside (mask); surely inside, then hits (hit)
            uint64_t mask = 0, hit = 0;
            int k = k0;
            for (int t = 0; t < wlen; t++) {
                const float Rk = sR[k][b];
                float dx = dxij + Rk * scs[0][k];
                float dy = dyij + Rk * scs[1][k];
                if (wrap2) {
                    dx = min_image(dx, Lxf);
                    dy = min_image(dy, Lyf);
                }
                const float d2 = dx * dx + dy * dy;
                const uint64_t bit = 1ull << k;
                if (d2 <= lim2) mask |= bit;
                if (d2 < rin2) hit |= bit;
                if (++k == NTH_M) k = 0;
            }
            mask &= ~hit;
            if (mask) {
                float lj[NC];
#pragma unroll
                for (int c = 0; c < NC; c++) lj[c] = labj[c];
                while (mask) {
                    const int kk = __builtin_ctzll(mask);
                    mask &= mask - 1;
                    const float Rk = sR[kk][b];
                    float dx = dxij + Rk * scs[0][kk];
                    float dy = dyij + Rk * scs[1][kk];
                    if (wrap2) {
                        dx = min_image(dx, Lxf);
                        dy = min_image(dy, Lyf);
                    }
                    const float d2 = fmaxf(dx * dx + dy * dy, tiny_d2<float>());
                    const float rs = f_rsqrt(d2);
                    const float dist = d2 * rs;
                    const float cp = dx * rs, sp = dy * rs;
                    if (r0j * contour_horner<float, K>(lj, cp, sp) - dist >= 0.f) hit |= 1ull << kk;
                }
            }
#pragma unroll
            for (int c = 0; c < NCLS; c++) {
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
    // the four lanes of a body: add the other groups' counters
#pragma unroll
    for (int c = 0; c < NCLS; c++) {
#pragma unroll
        for (int o = 16; o < 64; o <<= 1) {
            uint64_t other[NPLANE];
#pragma unroll
            for (int q = 0; q < NPLANE; q++) other[q] = shfl_xor64(pl[c][q], o);
            planes_add(pl[c], other);
        }
    }

    // the lane's 10 nodes of each counter plane at bits 0..9
    uint32_t plh[NCLS][NPLANE];
#pragma unroll
    for (int c = 0; c < NCLS; c++)
#pragma unroll
        for (int q = 0; q < NPLANE; q++) plh[c][q] = (uint32_t)(pl[c][q] >> (g * NQ4));

    // ---- pass 2: synthesis again; the contour gradients of the lane's nodes
    // replace R, R', R'' in their register slots (scaled as in k_forces_mfma)
    synth4<K, TT, BF>(lab_row, live, g, lane, asyn, acc);
    float fx = 0.f, fy = 0.f, tt = 0.f;
    const float sc = live ? PULL_SCALE : 0.f;
    const float Wps = Wp * sc, Wbs = Wb * sc * inv_u * inv_u;
    const float pf = p.dphi * u * sc;
    float Jpf[NCLS];
#pragma unroll
    for (int c = 0; c < NCLS; c++) Jpf[c] = p.Jcls[c] * pf;
    bool hi = false;
#pragma unroll
    for (int c = 0; c < NCLS; c++)
#pragma unroll
        for (int q = 2; q < NPLANE; q++) hi |= plh[c][q] != 0;
    const bool hi_any = __ballot(hi) != 0;
#pragma unroll
    for (int j = 0; j < NQ4; j++) {
        const int k = g * NQ4 + j;
        const float a = SLOT4(j), ad = SLOT4(NQ4 + j);
        const float rs = f_rsqrt(a * a + ad * ad);
        float GR = Wps * (a * rs), GD = Wps * (ad * rs), GDD = 0.f;
        if constexpr (T == 4) {
            float j0, j1, j2;
            term_J<float>(FCC_TERM_BENDING, a, ad, SLOT4(2 * NQ4 + j), j0, j1, j2);
            GR += Wbs * j0;
            GD += Wbs * j1;
            GDD = Wbs * j2;
        }
        if (contact) {
            float pp = 0.f;
#pragma unroll
            for (int c = 0; c < NCLS; c++) {
                uint32_t n = __builtin_amdgcn_ubfe(plh[c][0], (uint32_t)j, 1u) |
                             (__builtin_amdgcn_ubfe(plh[c][1], (uint32_t)j, 1u) << 1);
                if (hi_any) {
#pragma unroll
                    for (int qq = 2; qq < NPLANE; qq++) n |= __builtin_amdgcn_ubfe(plh[c][qq], (uint32_t)j, 1u) << qq;
                }
                pp += Jpf[c] * (float)n;
            }
            const float c = scs[0][k], s = scs[1][k];
            GR += pp * a;
            fx += pp * (a * c + ad * s);
            fy += pp * (a * s - ad * c);
            tt += (-pp) * a * ad;
        }
        SLOT4(j) = GR;
        SLOT4(NQ4 + j) = GD;
        if constexpr (T == 4) SLOT4(2 * NQ4 + j) = GDD;
    }
    // slots past the lane's quantities (s ≥ 20 without Bending, s ≥ 30 with
    // it) hold the synthesis of zero basis rows, i.e. zeros: they add nothing
    if (debug && live) {
        // contour gradients and contact density at the lane's nodes (tests)
#pragma unroll
        for (int j = 0; j < NQ4; j++) {
            const int64_t o = ii * NTH_M + g * NQ4 + j;
            d.GR[o] = SLOT4(j) * (1.f / PULL_SCALE);
            d.GD[o] = SLOT4(NQ4 + j) * (1.f / PULL_SCALE);
            d.GDD[o] = T == 4 ? SLOT4(2 * NQ4 + j) * (1.f / PULL_SCALE) : 0.f;
            if (contact) {
                float gg = 0.f;
#pragma unroll
                for (int c = 0; c < NCLS; c++) {
                    uint32_t n = 0;
#pragma unroll
                    for (int qq = 0; qq < NPLANE; qq++) n |= __builtin_amdgcn_ubfe(plh[c][qq], (uint32_t)j, 1u) << qq;
                    gg += p.Jcls[c] * (float)n;
                }
                d.P[o] = gg;
            }
        }
    }
    fx = fx + __shfl_xor(fx, 16);
    fy = fy + __shfl_xor(fy, 16);
    tt = tt + __shfl_xor(tt, 16);
    fx = (fx + __shfl_xor(fx, 32)) * (1.f / PULL_SCALE);
    fy = (fy + __shfl_xor(fy, 32)) * (1.f / PULL_SCALE);
    tt = (tt + __shfl_xor(tt, 32)) * (u * (1.f / PULL_SCALE));

    // ---- pullback on MFMA: glabᵀ (bodies x coefs) = Gᵀ·S over the TT tiles
    f4v_t gacc = {};
#pragma unroll
    for (int t = 0; t < TT; t++) {
        float x[4];
#pragma unroll
        for (int r = 0; r < 4; r++) x[r] = acc[t][r];
        v_t gh, gl;
        MT::split(x, gh, gl);
        const v_t sh = load_op4<BF>(bpull + ((t * 2 + 0) * 64 + lane) * 4);
        const v_t sl = load_op4<BF>(bpull + ((t * 2 + 1) * 64 + lane) * 4);
        gacc = MT::mfma(gh, sh, gacc);
        gacc = MT::mfma(gl, sh, gacc);
        gacc = MT::mfma(gh, sl, gacc);
    }
#undef SLOT4
    // lane l: coefficient l % 16 of bodies 4g..4g+3 of the tile
    const int n = lane & 15;
    if (n < NC) {
        const int64_t body0 = (int64_t)blk * 16 + 4 * g;
        float *base = d.glab + body0 * LDC_M + n;
#pragma unroll
        for (int r = 0; r < 4; r++)
            if (body0 + r < p.N) base[r * LDC_M] = gacc[r] * (1.f / PULL_SCALE);
    }
    if (live && g == 0) {
        d.gx[ii] = fx;
        d.gy[ii] = fy;
        d.gth[ii] = tt;
        d.Ebody[ii] = E;
        if (d.th0) d.th0[ii] = d.th[ii];  // parent drag: θ at the start of the step
    }
    if constexpr (NW > 1) {
        // the same rows into the wave's LDS slice (the contact pass has read
        // its contour slots): body r of the wave at rows[r·W], its force and
        // torque after the 16 rows
        float *rows = &sR[0][0];
        if (n < NC) {
#pragma unroll
            for (int r = 0; r < 4; r++) rows[(4 * g + r) * W + n] = gacc[r] * (1.f / PULL_SCALE);
        }
        if (g == 0) {
            rows[16 * W + 3 * b] = fx;
            rows[16 * W + 3 * b + 1] = fy;
            rows[16 * W + 3 * b + 2] = tt;
        }
        __syncthreads();
        if (w != 0) return;
        fused_update<float, K, true>(p, d, &sRw[0][0][0], NTH_M * 16, 16, wg * 64, lane, q, clk, use_status == 2);
    }
}

template <int K, int T, int NCLS, bool BF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_forces_mfma4(const Params<float> p, const Ptrs<float> d,
                                                     const uint16_t *asyn, const uint16_t *bpull, int use_status,
                                                     int debug) {
    k_forces_mfma4_impl<K, T, NCLS, BF, 1>(p, d, asyn, bpull, use_status, debug, xcd_tile(blockIdx.x, gridDim.x));
}

// One-kernel step of the mixed path with four lanes per body (four waves, 64
// bodies per workgroup) reading the buffers of parity q; use_status 3, or 4
// for timing series.
template <int K, int T, int NCLS, bool BF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_step_mfma4(const Params<float> p, const Ptrs<float> d,
                                                     const uint16_t *asyn, const uint16_t *bpull, int use_status,
                                                     int q) {
    k_forces_mfma4_impl<K, T, NCLS, BF, 4>(p, d, asyn, bpull, use_status, 0, xcd_tile(blockIdx.x, gridDim.x), q);
}

// Ensemble form: tiles of 64 bodies, each replica's buffers by the parity
// k_gate left in its status (step_buffers).
template <int K, int T, int NCLS, bool BF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_step_mfma4_ens(const Params<float> *ps, const Ptrs<float> *ds,
                                                         const Tile *tiles, const uint16_t *asyn,
                                                         const uint16_t *bpull) {
    const Tile tl = load_uniform(tiles, blockIdx.x);
    if (tl.rep < 0) return;
    const FCC_CONST Params<float> &p = as_const(ps)[tl.rep];
    const Ptrs<float> d = step_buffers(load_uniform(ds, tl.rep));
    k_forces_mfma4_impl<K, T, NCLS, BF, 4>(p, d, asyn, bpull, 2, 0, tl.blk);
}

template <int K, int T, int NCLS, bool BF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_forces_mfma4_ens(const Params<float> *ps, const Ptrs<float> *ds,
                                                         const Tile *tiles, const uint16_t *asyn,
                                                         const uint16_t *bpull, int use_status, int debug) {
    const Tile tl = load_uniform(tiles, blockIdx.x);
    if (tl.rep < 0) return;
    const FCC_CONST Params<float> &p = as_const(ps)[tl.rep];
    const FCC_CONST Ptrs<float> &d = as_const(ds)[tl.rep];
    k_forces_mfma4_impl<K, T, NCLS, BF, 1>(p, d, asyn, bpull, use_status, debug, tl.blk);
}

}  // namespace fcc
