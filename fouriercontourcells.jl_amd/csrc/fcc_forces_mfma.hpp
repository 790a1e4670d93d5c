// fcc_forces_mfma.hpp — forces kernel of the mixed-precision path: the contour
// synthesis [R; R'; R''] = S·lab and the pullback glab = Sᵀ·G run on the
// gfx950 matrix cores (v_mfma_f32_32x32x16_{f16,bf16}) with a hi+lo split of
// both operands and FP32 accumulation; body terms and contact run in FP32.
//
// Layout: 32 bodies per wave, two lanes per body. Lane l works on body
// b = l % 32 of the wave and half h = l / 32 of its contour (nodes 20h..20h+19
// for nθ = 40). The 32x32 MFMA output holds, in lane l, rows
// 8(r/4) + 4h + r%4 (r = 0..15) of column l % 32 of each 32-row tile; the basis
// rows are permuted on the host so that the slots σ = 16t + r of lane half h
// are R (σ < 20), R' (σ < 40) and R'' (σ < 60, with Bending) at its nodes.
// Only R goes to LDS ([node][body], 5 KB per wave) for the contact pass,
// which reads the whole contour of the body; after the contact pass the
// synthesis is run a second time (a few MFMAs), and the contour gradients
// replace R, R', R'' in the same registers, where the pullback takes its A
// operand (the row slots of a K-chunk are exactly the lane's 8 rows of that
// chunk). The small LDS footprint (5.4 KB per wave) leaves the occupancy to
// the VGPR count.
//
// hi+lo: x = hi + lo with hi = rn(x), lo = rn(x - hi) in the 16-bit type;
// A·B ≈ Ah·Bh + Ah·Bl + Al·Bh. lab is pre-scaled by 2^10 and G by 2^8 (powers
// of two, exact) so FP16 parts stay in the normal range.
#pragma once

#include "fcc_forces_fast.hpp"

namespace fcc {

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef __bf16 b8_t __attribute__((ext_vector_type(8)));
typedef float f16v_t __attribute__((ext_vector_type(16)));

constexpr float SYN_SCALE = 1024.f;  // lab pre-scale
constexpr float PULL_SCALE = 256.f;  // contour-gradient pre-scale
constexpr int NTH_M = 40;            // nodes handled by the MFMA kernel
constexpr int NH = NTH_M / 2;        // nodes per lane half
constexpr int LDC_M = 16;            // row stride of the coefficient rows (ncoef ≤ 15)

template <bool BF>
struct MfmaType;
template <>
struct MfmaType<false> {
    typedef _Float16 s_t;
    typedef h8_t v_t;
    static __device__ __forceinline__ f16v_t mfma(v_t a, v_t b, f16v_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct MfmaType<true> {
    typedef __bf16 s_t;
    typedef b8_t v_t;
    static __device__ __forceinline__ f16v_t mfma(v_t a, v_t b, f16v_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

// Each value is pinned in a register as the rounded FP32 result of the
// operation that produced it (an empty asm with a read-write operand) before
// it is split: otherwise the compiler may fold the producing multiply into
// the conversion (v_fma_mix, one rounding to 16 bits instead of two) in some
// kernels and not in others, and kernels that share this code would round
// differently.
template <bool BF>
__device__ __forceinline__ void split8(const float (&x)[8], typename MfmaType<BF>::v_t &hi,
                                       typename MfmaType<BF>::v_t &lo) {
    typedef typename MfmaType<BF>::s_t s_t;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        float v = x[s];
        asm volatile("" : "+v"(v));
        const s_t h = (s_t)v;
        hi[s] = h;
        lo[s] = (s_t)(v - (float)h);
    }
}

template <bool BF>
__device__ __forceinline__ typename MfmaType<BF>::v_t load_op(const uint16_t *p) {
    typename MfmaType<BF>::v_t v;
    const uint4 q = *reinterpret_cast<const uint4 *>(p);
    __builtin_memcpy(&v, &q, 16);
    return v;
}

// Synthesis on MFMA: lane (b, h) supplies lab coefficients 8h..8h+7 of body
// b (pre-scaled by SYN_SCALE); acc[t] receives tile t of S·(SYN_SCALE·lab)
// (slots σ = 16t + r). The contour is R = u·acc with u = R0 / SYN_SCALE; the
// consumers fold u into per-body factors instead of rescaling every slot.
template <int K, int T, bool BF>
__device__ __forceinline__ void synth_tiles(const float *lab_row, bool live, int h, int lane,
                                            const uint16_t *asyn, f16v_t (&acc)[T]) {
    typedef MfmaType<BF> MT;
    typedef typename MT::v_t v_t;
    constexpr int NC = 2 * K + 1;
    float x[8];
    const float4 *row = reinterpret_cast<const float4 *>(lab_row + 8 * h);
    const float4 u0 = row[0], u1 = row[1];
    x[0] = u0.x; x[1] = u0.y; x[2] = u0.z; x[3] = u0.w;
    x[4] = u1.x; x[5] = u1.y; x[6] = u1.z; x[7] = u1.w;
#pragma unroll
    for (int s = 0; s < 8; s++) x[s] = (live && 8 * h + s < NC) ? x[s] * SYN_SCALE : 0.f;
    v_t bh, bl;
    split8<BF>(x, bh, bl);
#pragma unroll
    for (int t = 0; t < T; t++) {
        const v_t ah = load_op<BF>(asyn + ((t * 2 + 0) * 64 + lane) * 8);
        const v_t al = load_op<BF>(asyn + ((t * 2 + 1) * 64 + lane) * 8);
        f16v_t c = {};
        c = MT::mfma(ah, bh, c);
        c = MT::mfma(ah, bl, c);
        c = MT::mfma(al, bh, c);
        acc[t] = c;
    }
}

// Operand tables (built on the host, 16-bit values):
//  asyn[(t * 2 + part) * 64 + lane][8]   synthesis A (S rows), part 0 hi / 1 lo
//  bpull[(c * 2 + v) * 64 + lane][8]     pullback B, v 0 = [Sh | Sl], 1 = [Sh | 0]
//
// NW waves per workgroup, wave w on the 32 bodies of tile wg·NW + w. With
// NW = 2 the workgroup is a one-kernel step: the waves leave their bodies'
// pullback rows and centre force / torque in their LDS slices (the contour
// slots are dead after the contact pass), and wave 0 updates the 64 bodies
// (fused_update).
template <int K, int T, int NCLS, bool BF, int NW, typename PP, typename DP>
__device__ __forceinline__ void k_forces_mfma_impl(const PP &p, const DP &d,
                                                    const uint16_t *asyn, const uint16_t *bpull, int use_status,
                                                    int debug, int64_t wg, int q = 0) {
    typedef MfmaType<BF> MT;
    typedef typename MT::v_t v_t;
    constexpr int NC = 2 * K + 1;
    constexpr int W = row_width<NC, float>();
    static_assert(NW == 1 || NW == 2, "one or two waves per workgroup");
    static_assert(32 * W + 3 * 32 <= NTH_M * 32, "pullback rows fit the wave's LDS slice");
    __shared__ float sRw[NW][NTH_M][32];  // R of each wave's 32 bodies for the contact pass
    __shared__ float scs[2][NTH_M];
    const int lane = threadIdx.x & 63, w = NW == 1 ? 0 : (int)(threadIdx.x >> 6);
    float(*sR)[32] = sRw[w];
    const int b = lane & 31, h = lane >> 5;
    for (int k = threadIdx.x; k < NTH_M; k += 64 * NW) {
        scs[0][k] = d.cs[2 * k];
        scs[1][k] = d.cs[2 * k + 1];
    }
    const int64_t blk = wg * NW + w;
    const int64_t i = (int64_t)blk * 32 + b;
    int64_t clk = 0;  // one-kernel step: its clock
    if constexpr (NW == 2) {
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
    const float *lab_row = d.lab + ii * LDC_M;

    // ---- synthesis on MFMA
    f16v_t acc[T];
    synth_tiles<K, T, BF>(lab_row, live, h, lane, asyn, acc);
    // slot σ = 16t + r → acc[t][r]; R/u at σ = j, R'/u at 20 + j, R''/u at 40 + j
#define SLOT(sig) acc[(sig) >> 4][(sig)&15]

    const int ti = live ? d.type[ii] : 0;
    const uint32_t tbit = 1u << ti;
    const bool contact = p.ncontact != 0;

    // ---- pass 1: perimeter and bending integrals over both halves, from the
    // unscaled slots: √(R² + R'²) is homogeneous of degree 1 in (R, R', R''),
    // the bending integrand of degree -1
    float Qp = 0.f, Qb = 0.f;
#pragma unroll
    for (int j = 0; j < NH; j++) {
        const float a = SLOT(j), ad = SLOT(NH + j);
        Qp += f_sqrt(a * a + ad * ad);
        if constexpr (T == 4) Qb += term_q<float>(FCC_TERM_BENDING, a, ad, SLOT(2 * NH + j));
        sR[h * NH + j][b] = a * u;
    }
    Qp = (Qp + __shfl_xor(Qp, 32)) * u;
    const float inv_u = SYN_SCALE / r0;
    if constexpr (T == 4) Qb = (Qb + __shfl_xor(Qb, 32)) * inv_u;
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

    // ---- contact: lane half h takes neighbour slots q ≡ h (mod 2), full circle
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
        int jn = h < cnt ? nrow[(int64_t)h * p.N] : 0;
        for (int q = h; q < cnt; q += 2) {
            const int j = jn;
            const double xj = d.x[j], yj = d.y[j];
            const float rmaxj = d.rmax[j], rminj = d.rmin[j], r0j = d.R0[j];
            float labj[W];
            row_load<float, W>(labj, d.lab + (int64_t)j * LDC_M);
            if (q + 2 < cnt) jn = nrow[(int64_t)(q + 2) * p.N];
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
    // both lanes of a body: add the partner's counters
#pragma unroll
    for (int c = 0; c < NCLS; c++) {
        uint64_t other[NPLANE];
#pragma unroll
        for (int q = 0; q < NPLANE; q++) other[q] = shfl_xor64(pl[c][q], 32);
        planes_add(pl[c], other);
    }

    // the lane's 20 nodes of each counter plane at bits 0..19
    uint32_t plh[NCLS][NPLANE];
#pragma unroll
    for (int c = 0; c < NCLS; c++)
#pragma unroll
        for (int q = 0; q < NPLANE; q++) plh[c][q] = (uint32_t)(pl[c][q] >> (h * NH));

    // ---- pass 2: synthesis again; the contour gradients of the lane's nodes
    // replace R, R', R'' in their register slots
    synth_tiles<K, T, BF>(lab_row, live, h, lane, asyn, acc);
    // Scaled gradients GR·sc, GD·sc, GDD·sc (sc = PULL_SCALE, a power of two)
    // from the unscaled slots a = R/u: R·rsqrt(R² + R'²) is scale free, the
    // bending gradients are of degree -2, the contact terms take u into the
    // per-body factor pf; fx, fy, tt carry sc (and tt one more u) until the end.
    float fx = 0.f, fy = 0.f, tt = 0.f;
    const float sc = live ? PULL_SCALE : 0.f;
    const float Wps = Wp * sc, Wbs = Wb * sc * inv_u * inv_u;
    const float pf = p.dphi * u * sc;
    float Jpf[NCLS];
#pragma unroll
    for (int c = 0; c < NCLS; c++) Jpf[c] = p.Jcls[c] * pf;
    // Planes 2.. hold counts of 4 or more (a node inside four neighbours):
    // they are read only when some lane of the wave has one.
    bool hi = false;
#pragma unroll
    for (int c = 0; c < NCLS; c++)
#pragma unroll
        for (int q = 2; q < NPLANE; q++) hi |= plh[c][q] != 0;
    const bool hi_any = __ballot(hi) != 0;
#pragma unroll
    for (int j = 0; j < NH; j++) {
        const int k = h * NH + j;
        const float a = SLOT(j), ad = SLOT(NH + j);
        const float rs = f_rsqrt(a * a + ad * ad);
        float GR = Wps * (a * rs), GD = Wps * (ad * rs), GDD = 0.f;
        if constexpr (T == 4) {
            float j0, j1, j2;
            term_J<float>(FCC_TERM_BENDING, a, ad, SLOT(2 * NH + j), j0, j1, j2);
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
        SLOT(j) = GR;
        SLOT(NH + j) = GD;
        if constexpr (T == 4) SLOT(2 * NH + j) = GDD;
    }
    if (debug && live) {
        // contour gradients and contact density at the lane's nodes (tests)
#pragma unroll
        for (int j = 0; j < NH; j++) {
            const int64_t o = ii * NTH_M + h * NH + j;
            d.GR[o] = SLOT(j) * (1.f / PULL_SCALE);
            d.GD[o] = SLOT(NH + j) * (1.f / PULL_SCALE);
            d.GDD[o] = T == 4 ? SLOT(2 * NH + j) * (1.f / PULL_SCALE) : 0.f;
            if (contact) {
                float g = 0.f;
#pragma unroll
                for (int c = 0; c < NCLS; c++) {
                    uint32_t n = 0;
#pragma unroll
                    for (int qq = 0; qq < NPLANE; qq++) n |= __builtin_amdgcn_ubfe(plh[c][qq], (uint32_t)j, 1u) << qq;
                    g += p.Jcls[c] * (float)n;
                }
                d.P[o] = g;
            }
        }
    }
    fx = (fx + __shfl_xor(fx, 32)) * (1.f / PULL_SCALE);
    fy = (fy + __shfl_xor(fy, 32)) * (1.f / PULL_SCALE);
    tt = (tt + __shfl_xor(tt, 32)) * (u * (1.f / PULL_SCALE));

    // ---- pullback on MFMA: glabᵀ (bodies x coefs) = Gᵀ·S over 2T K-chunks of 16 rows
    f16v_t gacc = {};
#pragma unroll
    for (int c = 0; c < 2 * T; c++) {
        float x[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int sig = 16 * (c >> 1) + 8 * (c & 1) + s;  // slot σ = 16t + r
            x[s] = sig < (T == 4 ? 3 : 2) * NH ? SLOT(sig) : 0.f;
        }
        v_t ah, al;
        split8<BF>(x, ah, al);
        const v_t b1 = load_op<BF>(bpull + ((c * 2 + 0) * 64 + lane) * 8);
        const v_t b2 = load_op<BF>(bpull + ((c * 2 + 1) * 64 + lane) * 8);
        gacc = MT::mfma(ah, b1, gacc);
        gacc = MT::mfma(al, b2, gacc);
    }
#undef SLOT
    // column n of lane l: n < 16 hi·hi + lo·hi of coefficient n, n ≥ 16 hi·lo of n - 16
#pragma unroll
    for (int r = 0; r < 16; r++) gacc[r] += __shfl_down(gacc[r], 16);
    const int n = lane & 31;
    if (n < NC) {
        const int64_t body0 = (int64_t)blk * 32 + 4 * h;
        float *base = d.glab + body0 * LDC_M + n;
        if ((int64_t)blk * 32 + 32 <= p.N) {
#pragma unroll
            for (int r = 0; r < 16; r++) base[(8 * (r >> 2) + (r & 3)) * LDC_M] = gacc[r] * (1.f / PULL_SCALE);
        } else {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int rr = 8 * (r >> 2) + (r & 3);
                if (body0 + rr < p.N) base[rr * LDC_M] = gacc[r] * (1.f / PULL_SCALE);
            }
        }
    }
    if (live && h == 0) {
        d.gx[ii] = fx;
        d.gy[ii] = fy;
        d.gth[ii] = tt;
        d.Ebody[ii] = E;
        if (d.th0) d.th0[ii] = d.th[ii];  // parent drag: θ at the start of the step
    }
    if constexpr (NW == 2) {
        // the same rows into the wave's LDS slice: body rr of the wave at
        // rows[rr·W], its force and torque after the 32 rows
        float *rows = &sR[0][0];
        if (n < NC) {
#pragma unroll
            for (int r = 0; r < 16; r++) rows[(4 * h + 8 * (r >> 2) + (r & 3)) * W + n] = gacc[r] * (1.f / PULL_SCALE);
        }
        if (h == 0) {
            rows[32 * W + 3 * b] = fx;
            rows[32 * W + 3 * b + 1] = fy;
            rows[32 * W + 3 * b + 2] = tt;
        }
        __syncthreads();
        if (w != 0) return;
        fused_update<float, K, true>(p, d, &sRw[0][0][0], NTH_M * 32, 32, wg * 64, lane, q, clk, use_status == 2);
    }
}

template <int K, int T, int NCLS, bool BF>
__global__ __launch_bounds__(64) void k_forces_mfma(const Params<float> p, const Ptrs<float> d,
                                                    const uint16_t *asyn, const uint16_t *bpull, int use_status,
                                                    int debug) {
    k_forces_mfma_impl<K, T, NCLS, BF, 1>(p, d, asyn, bpull, use_status, debug, xcd_tile(blockIdx.x, gridDim.x));
}

// One-kernel step of the mixed path (two waves, 64 bodies per workgroup)
// reading the buffers of parity q; use_status 3, or 4 for timing series.
template <int K, int T, int NCLS, bool BF>
__global__ __launch_bounds__(128) void k_step_mfma(const Params<float> p, const Ptrs<float> d,
                                                   const uint16_t *asyn, const uint16_t *bpull, int use_status,
                                                   int q) {
    k_forces_mfma_impl<K, T, NCLS, BF, 2>(p, d, asyn, bpull, use_status, 0, xcd_tile(blockIdx.x, gridDim.x), q);
}

// Ensemble form of the one-kernel step: tiles of 64 bodies, each replica's
// buffers by the parity k_gate left in its status (step_buffers).
template <int K, int T, int NCLS, bool BF>
__global__ __launch_bounds__(128) void k_step_mfma_ens(const Params<float> *ps, const Ptrs<float> *ds,
                                                       const Tile *tiles, const uint16_t *asyn,
                                                       const uint16_t *bpull) {
    const Tile tl = load_uniform(tiles, blockIdx.x);
    if (tl.rep < 0) return;
    const FCC_CONST Params<float> &p = as_const(ps)[tl.rep];
    const Ptrs<float> d = step_buffers(load_uniform(ds, tl.rep));
    k_forces_mfma_impl<K, T, NCLS, BF, 2>(p, d, asyn, bpull, 2, 0, tl.blk);
}

template <int K, int T, int NCLS, bool BF>
__global__ __launch_bounds__(64) void k_forces_mfma_ens(const Params<float> *ps, const Ptrs<float> *ds,
                                                        const Tile *tiles, const uint16_t *asyn,
                                                        const uint16_t *bpull, int use_status, int debug) {
    const Tile tl = load_uniform(tiles, blockIdx.x);
    if (tl.rep < 0) return;
    const FCC_CONST Params<float> &p = as_const(ps)[tl.rep];
    const FCC_CONST Ptrs<float> &d = as_const(ds)[tl.rep];
    k_forces_mfma_impl<K, T, NCLS, BF, 1>(p, d, asyn, bpull, use_status, debug, tl.blk);
}

}  // namespace fcc
