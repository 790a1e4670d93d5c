// fcc_observe.hpp — long-run observables accumulated on the device.
//
// The reference computes them on the host from full copies of the state
// (src/observables.jl:6-29): shape index p = perimeter / sqrt(area) per body,
// and the unwrapped positions whose mean squared displacement the examples
// and the long-run statistics follow. A sample here is two kernels on the
// simulation's stream and no host transfer:
//   k_observe        one thread per body: p (the same contour sums as
//                    k_geometry), |r_unwrapped - r_origin|^2, the body's bin of
//                    the per-type shape-index histogram (LDS counters, one
//                    global atomic per non-empty bin and workgroup), and
//                    per-type partial sums {bodies, sum p, sum p^2, sum |dr|^2}
//                    per workgroup;
//   k_observe_finish one workgroup: the partial sums in a fixed order (the
//                    sample's means do not depend on scheduling) into row s
//                    of the sample series, with the device clock.
#pragma once

#include "fcc_kernels.hpp"

namespace fcc {

constexpr int OBS_BLOCK = 256;
constexpr int OBS_NV = 4;  // per type: bodies, sum p, sum p^2, sum |dr|^2

struct ObsArgs {
    const double *x0, *y0;     // unwrapped positions at the start of the observation
    double pmin, scale;        // bin of p: 1 + floor((p - pmin) * scale) in 1..nbins
    int32_t nbins, ntypes;     // bins 0 and nbins+1: below pmin, at or above pmax (or NaN)
    double *part;              // [workgroups][ntypes][OBS_NV]
    unsigned long long *hist;  // [ntypes][nbins + 2], accumulated over samples
};

// Unwrapped centre positions (src/observables.jl:24-29), evaluated as there:
// x + imgx * Lx without contraction.
__global__ __launch_bounds__(256) void k_unwrapped(int64_t N, const double *x, const double *y, const int32_t *imgx,
                                                   const int32_t *imgy, double Lx, double Ly, double *ux,
                                                   double *uy) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    ux[i] = x[i] + (double)imgx[i] * Lx;
    uy[i] = y[i] + (double)imgy[i] * Ly;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

template <typename T, int KM>
__global__ __launch_bounds__(OBS_BLOCK) void k_observe(const Params<T> p, const Ptrs<T> d, const ObsArgs o) {
    extern __shared__ uint32_t hcount[];  // [ntypes][nbins + 2]
    __shared__ double wpart[OBS_BLOCK / 64][FCC_MAX_TYPES * OBS_NV];
    const int nh = o.ntypes * (o.nbins + 2);
    for (int q = threadIdx.x; q < nh; q += OBS_BLOCK) hcount[q] = 0u;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * OBS_BLOCK + threadIdx.x;
    int ti = -1;
    double pv = 0.0, d2 = 0.0;
    if (i < p.N) {
        ti = d.type[i];
        T per, area;
        contour_sums<T, KM>(p, d, i, nullptr, nullptr, nullptr, per, area);
        const T pt = per / sqrt(area);
        pv = (double)pt;
        {
#pragma clang fp contract(off)
            const double dx = (d.x[i] + (double)d.imgx[i] * p.Lx) - o.x0[i];
            const double dy = (d.y[i] + (double)d.imgy[i] * p.Ly) - o.y0[i];
            d2 = dx * dx + dy * dy;
            int bin;
            if (pv < o.pmin)
                bin = 0;
            else {
                const double u = floor((pv - o.pmin) * o.scale);
                bin = u < (double)o.nbins ? 1 + (int)u : o.nbins + 1;  // NaN compares false: last bin
            }
            atomicAdd(&hcount[ti * (o.nbins + 2) + bin], 1u);
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int t = 0; t < o.ntypes; t++) {
        const bool mine = ti == t;
        const double c = wave_sum(mine ? 1.0 : 0.0);
        const double sp = wave_sum(mine ? pv : 0.0);
        const double sp2 = wave_sum(mine ? pv * pv : 0.0);
        const double sd = wave_sum(mine ? d2 : 0.0);
        if (lane == 0) {
            wpart[w][t * OBS_NV] = c;
            wpart[w][t * OBS_NV + 1] = sp;
            wpart[w][t * OBS_NV + 2] = sp2;
            wpart[w][t * OBS_NV + 3] = sd;
        }
    }
    __syncthreads();
    const int nv = o.ntypes * OBS_NV;
    if ((int)threadIdx.x < nv) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < OBS_BLOCK / 64; q++) s += wpart[q][threadIdx.x];
        o.part[(int64_t)blockIdx.x * nv + threadIdx.x] = s;
    }
    for (int q = threadIdx.x; q < nh; q += OBS_BLOCK)
        if (hcount[q]) atomicAdd(&o.hist[q], (unsigned long long)hcount[q]);
}

// Row s of the sample series: per type the mean |dr|^2, mean p and mean p^2
// over the bodies of that type (NaN for a type without bodies).
__global__ __launch_bounds__(OBS_BLOCK) void k_observe_finish(const double *part, int64_t nblk, int32_t ntypes,
                                                              int64_t s, const DevStatus *st, double *msd,
                                                              double *pmean, double *p2mean, int64_t *clock) {
    __shared__ double wsum[OBS_BLOCK / 64][FCC_MAX_TYPES * OBS_NV];
    const int nv = ntypes * OBS_NV;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int v = 0; v < nv; v++) {
        double a = 0.0;
        for (int64_t b = threadIdx.x; b < nblk; b += OBS_BLOCK) a += part[b * nv + v];
        a = wave_sum(a);
        if (lane == 0) wsum[w][v] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int t = 0; t < ntypes; t++) {
            double tot[OBS_NV];
            for (int c = 0; c < OBS_NV; c++) {
                double a = 0.0;
                for (int q = 0; q < OBS_BLOCK / 64; q++) a += wsum[q][t * OBS_NV + c];
                tot[c] = a;
            }
            const double n = tot[0];
            msd[s * ntypes + t] = n > 0.0 ? tot[3] / n : __builtin_nan("");
            pmean[s * ntypes + t] = n > 0.0 ? tot[1] / n : __builtin_nan("");
            p2mean[s * ntypes + t] = n > 0.0 ? tot[2] / n : __builtin_nan("");
        }
        clock[s] = st->clock;
    }
}

// contours(sim) (src/observables.jl:36-44): contour nodes around the wrapped
// centres, X = x + R cos(phi_k), Y = y + R sin(phi_k) (n_theta x N, column-major).
template <typename T, int KM>
__global__ __launch_bounds__(256) void k_contours(const Params<T> p, const Ptrs<T> d, double *X, double *Y) {
    constexpr int NCM = 2 * KM + 1;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= p.N) return;
    const FCC_CONST T *S = as_const(d.S);
    T l[NCM];
    load_row<T, KM>(l, d.lab + i * p.ldc, p.K);
    const T r0 = d.R0[i];
    const double xi = d.x[i], yi = d.y[i];
    for (int k = 0; k < p.nth; k++) {
        T r, dr, ddr;
        synth<T, KM>(S, p.nth, k, l, p.K, false, r0, r, dr, ddr);
        X[i * p.nth + k] = xi + (double)r * (double)d.cs[2 * k];
        Y[i * p.nth + k] = yi + (double)r * (double)d.cs[2 * k + 1];
    }
}

}  // namespace fcc
