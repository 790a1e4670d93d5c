// fcc_forces_wide.hpp — gfx950 forces kernel for contours of more than 64
// nodes (up to FCC_MAX_THETA): one workgroup per body, node k on lane
// k mod WG, the contact density of every node in LDS.
//
// The per-node arithmetic is that of k_forces (body terms
// src/gradients.jl:70-128, Overlap / LinearRepulsion contact
// src/energies/contact.jl:58-82,106-134,149-187, pullback
// src/gradients.jl:374-386); only the layout differs. Every node belongs to
// one lane for the whole launch, so the contact density of a node is summed
// over the neighbour list in list order by that lane alone, with no barrier
// between neighbours. The sums over nodes (body-term integrals, pullback,
// force and torque) are per-lane partial sums combined by a fixed butterfly,
// so results do not depend on scheduling.
#pragma once

#include "fcc_kernels.hpp"

namespace fcc {

// Contours up to this many nodes take one wave per body, larger ones four.
constexpr int WIDE_WG64_THETA = 2048;

// R, R', R'' at node k, as synth computes them, for a node index that differs
// between lanes (vector loads of the basis rows).
template <typename T, int KM>
__device__ __forceinline__ void synth_lane(const T *S, int nth, int k, const T (&l)[2 * KM + 1], int K,
                                           bool need_dd, T r0, T &R, T &dR, T &ddR) {
    const T *C = S + (size_t)k * SB;
    const T *D = S + (size_t)(nth + k) * SB;
    const T *DD = S + (size_t)(2 * nth + k) * SB;
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

// Sum of v over the WG lanes of the workgroup; every lane receives the total.
// Within a wave an xor butterfly (the two lanes of each pair add the same two
// operands, so all lanes hold the same bits), then the wave totals in wave
// order through red (WG / 64 × NV values).
template <typename T, int WG, int NV>
__device__ __forceinline__ void wg_sum(T (&v)[NV], T *red) {
#pragma unroll
    for (int c = 0; c < NV; c++)
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) v[c] += __shfl_xor(v[c], o);
    if constexpr (WG > 64) {
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0)
#pragma unroll
            for (int c = 0; c < NV; c++) red[w * NV + c] = v[c];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NV; c++) {
            T s = red[c];
            for (int q = 1; q < WG / 64; q++) s += red[q * NV + c];
            v[c] = s;
        }
        __syncthreads();
    }
}

// Forces of body i (writes glab, gx, gy, gθ, Ebody and, with debug, P, GR,
// GD, GDD at every node). blk is the workgroup that opens the step for
// use_status 1 (workgroup 0).
template <typename T, int KM, int WG, typename PP, typename DP>
__device__ __forceinline__ void k_forces_wide_impl(const PP &p, const DP &d, int use_status, int debug,
                                                    int64_t i, int64_t blk) {
    constexpr int NCM = 2 * KM + 1;
    constexpr int NV = NCM + 3;
    extern __shared__ __align__(16) unsigned char wide_lds[];
    __shared__ T red[(WG / 64) * NV];
    T *sP = reinterpret_cast<T *>(wide_lds);  // contact density, nθ values
    const int tid = threadIdx.x;
    if (forces_frozen(use_status, d.st, blk, tid)) return;
    if (i >= p.N) return;
    const int nth = p.nth, K = p.K;
    const T *S = d.S;
    const T *cs = d.cs;
    T lab[NCM];
    load_row<T, KM>(lab, d.lab + i * p.ldc, K);
    const T r0 = d.R0[i];
    const int ti = d.type[i];
    const uint32_t tbit = 1u << ti;
    const bool need_dd = p.need_dd != 0;
    const bool contact = p.ncontact != 0;

    // pass 1: contour integrals Q of every body term
    T Q[MAXB];
#pragma unroll
    for (int b = 0; b < MAXB; b++) Q[b] = T(0);
    for (int k = tid; k < nth; k += WG) {
        T R, dR, ddR;
        synth_lane<T, KM>(S, nth, k, lab, K, need_dd, r0, R, dR, ddR);
#pragma unroll
        for (int b = 0; b < MAXB; b++)
            if (b < p.nbody) Q[b] += term_q<T>(p.bkind[b], R, dR, ddR);
        if (contact) sP[k] = T(0);
    }
    wg_sum<T, WG, MAXB>(Q, red);
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

    // contact density at every node, over the neighbour list in list order;
    // the nodes of the angular window that can reach the neighbour's bounding
    // disc are tested, each by the lane that owns it
    if (contact) {
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
            T lj[NCM];
            load_row<T, KM>(lj, d.lab + (int64_t)j * p.ldc, K);
            const T r0j = d.R0[j];
            const T Jv = p.Jc[pi == d.parent[j] ? 1 : 0][ti][d.type[j]];
            const T kr = p.krep[ti];
            // the window is the node run [k0, k0 + n1) and, when it wraps, [0, wlen - n1)
            const int n1 = min(wlen, nth - k0);
#pragma unroll
            for (int run = 0; run < 2; run++) {
                const int a = run ? 0 : k0;
                const int e = run ? wlen - n1 : k0 + n1;
                for (int k = a + ((tid - a) & (WG - 1)); k < e; k += WG) {
                    T Rk, dRk, ddRk;
                    synth_lane<T, KM>(S, nth, k, lab, K, false, r0, Rk, dRk, ddRk);
                    const T dx = min_image(dxij + Rk * cs[2 * k], Lxf);
                    const T dy = min_image(dyij + Rk * cs[2 * k + 1], Lyf);
                    if (dx * dx + dy * dy <= lim2) {
                        const T dist = sqrt(dx * dx + dy * dy);
                        const T cp = dist > T(0) ? dx / dist : T(1);
                        const T sp = dist > T(0) ? dy / dist : T(0);
                        const T dd = r0j * contour_dir<T, KM>(lj, K, cp, sp) - dist;
                        sP[k] += (dd >= T(0) ? Jv : T(0)) + (dd > T(0) ? kr * dd : T(0));
                    }
                }
            }
        }
    }

    // pass 2: contour gradients per node, contact force/torque, pullback
    T acc[NV];  // glab row, then fx, fy, gθ
#pragma unroll
    for (int c = 0; c < NV; c++) acc[c] = T(0);
    for (int k = tid; k < nth; k += WG) {
        T R, dR, ddR;
        synth_lane<T, KM>(S, nth, k, lab, K, need_dd, r0, R, dR, ddR);
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
            const T g = sP[k];
            const T pp = g * p.dphi;
            const T c = cs[2 * k], s = cs[2 * k + 1];
            GR += pp * R;
            acc[NCM] += pp * (R * c + dR * s);
            acc[NCM + 1] += pp * (R * s - dR * c);
            acc[NCM + 2] += (-pp) * R * dR;
            if (debug) d.P[i * nth + k] = g;
        }
        const T *C = S + (size_t)k * SB;
        const T *D = S + (size_t)(nth + k) * SB;
        const T *DD = S + (size_t)(2 * nth + k) * SB;
        acc[0] += C[0] * GR;
#pragma unroll
        for (int m = 1; m <= KM; m++) {
            if (m <= K) {
                T a = C[m] * GR + D[m] * GD;
                T b = C[SH + m] * GR + D[SH + m] * GD;
                if (need_dd) {
                    a += DD[m] * GDD;
                    b += DD[SH + m] * GDD;
                }
                acc[m] += a;
                acc[KM + m] += b;
            }
        }
        if (debug) {
            d.GR[i * nth + k] = GR;
            d.GD[i * nth + k] = GD;
            d.GDD[i * nth + k] = GDD;
        }
    }
    wg_sum<T, WG, NV>(acc, red);
    if (tid == 0) {
        T gl[NCM];
#pragma unroll
        for (int c = 0; c < NCM; c++) gl[c] = acc[c];
        store_row<T, KM>(gl, d.glab + i * p.ldc, K);
        d.gx[i] = acc[NCM];
        d.gy[i] = acc[NCM + 1];
        d.gth[i] = acc[NCM + 2];
        d.Ebody[i] = E;
        if (d.th0) d.th0[i] = d.th[i];  // parent drag: θ at the start of the step
    }
}

// One workgroup per body; dynamic LDS of nθ values of T.
template <typename T, int KM, int WG>
__global__ __launch_bounds__(WG) void k_forces_wide(const Params<T> p, const Ptrs<T> d, int use_status, int debug) {
    k_forces_wide_impl<T, KM, WG>(p, d, use_status, debug, blockIdx.x, blockIdx.x);
}

// Ensemble launch: tile b is body tiles[b].blk of replica tiles[b].rep.
template <typename T, int KM, int WG>
__global__ __launch_bounds__(WG) void k_forces_wide_ens(const Params<T> *ps, const Ptrs<T> *ds, const Tile *tiles,
                                                        int use_status, int debug) {
    const Tile tl = load_uniform(tiles, blockIdx.x);
    if (tl.rep < 0) return;
    const FCC_CONST Params<T> &p = as_const(ps)[tl.rep];
    const FCC_CONST Ptrs<T> &d = as_const(ds)[tl.rep];
    k_forces_wide_impl<T, KM, WG>(p, d, use_status, debug, tl.blk, tl.blk);
}

}  // namespace fcc
