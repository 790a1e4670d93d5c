// fcc_api.hip — host orchestration of the gfx950 time step and the C ABI of
// include/fcc.h (libfcc_hip.so).
//
// One simulation = one HIP stream and device buffers it owns. A step is two
// kernels (forces, update; three with parent drag), eager (fcc_step, one
// host read of the rebuild flag per step, like step! in src/simulation.jl:165-173)
// or replayed from a captured hipGraph in chunks (fcc_run, like
// run_steps! in ext/FourierContourCellsCUDAExt.jl:144-180).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fcc_forces_mfma.hpp"
#include "fcc_forces_mfma4.hpp"
#include "fcc_forces_wide.hpp"
#include "fcc_observe.hpp"
#include "../../include/fcc_bench.h"

using namespace fcc;

static thread_local char g_err[512];

static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) return fail(FCC_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

static inline unsigned blocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Kernel launch; given timing events (EB at the start, EE at the end of the
// kernel) the dispatch itself records them (hipExtLaunchKernelGGL), with no
// marker packets between kernels, so the interval is the kernel's duration.
#define FCC_LAUNCH(KERN, G, B, SH, S, EB, EE, ...)                                                    \
    do {                                                                                             \
        if ((EB) || (EE))                                                                            \
            hipExtLaunchKernelGGL(KERN, dim3(G), dim3(B), SH, S, EB, EE, 0, __VA_ARGS__);            \
        else                                                                                         \
            hipLaunchKernelGGL(KERN, dim3(G), dim3(B), SH, S, __VA_ARGS__);                          \
    } while (0)

constexpr int64_t FOUR_LANES_BELOW = 131072;

// Steps per long graph (fcc_run and ensemble runs replay it for all but the
// last few steps of a chunk).
constexpr int GRAPH_STEPS = 8;

// Graph capture is serialised (ext/FourierContourCellsCUDAExt.jl:80-84).
static std::mutex g_capture_lock;

// ------------------------------------------------------------- validation

// Simulation(...) checks (src/simulation.jl:71-118, src/basis.jl:36-38,
// src/neighbors.jl:123) and the native-kernel limits.
static int validate(const fcc_config *cfg, const fcc_bodies *b) {
    if (!cfg || !b) return fail(FCC_ERR_ARGUMENT, "null argument");
    if (cfg->n_bodies < 1) return fail(FCC_ERR_ARGUMENT, "need at least one body");
    if (cfg->n_modes < 1) return fail(FCC_ERR_ARGUMENT, "maxmode must be ≥ 2");
    if (cfg->n_theta <= 2 * (cfg->n_modes + 1)) return fail(FCC_ERR_ARGUMENT, "nθ must exceed 2·maxmode");
    if (cfg->skin < 0.0) return fail(FCC_ERR_ARGUMENT, "skin must be non-negative");
    if (cfg->n_types < 1 || cfg->n_types > FCC_MAX_TYPES)
        return fail(FCC_ERR_UNSUPPORTED, "1..%d body types are native", FCC_MAX_TYPES);
    if (cfg->n_terms < 0 || cfg->n_terms > FCC_MAX_TERMS)
        return fail(FCC_ERR_UNSUPPORTED, "at most %d energy terms", FCC_MAX_TERMS);
    if (cfg->n_modes > FCC_MAX_MODES) return fail(FCC_ERR_UNSUPPORTED, "at most %d shape modes are native", FCC_MAX_MODES);
    if (cfg->n_theta > FCC_MAX_THETA) return fail(FCC_ERR_UNSUPPORTED, "nθ > %d has no native kernel", FCC_MAX_THETA);
    if (cfg->n_bodies > (int64_t)INT32_MAX) return fail(FCC_ERR_UNSUPPORTED, "too many bodies");
    if (cfg->precision < FCC_FP64 || cfg->precision > FCC_MIXED_BF16) return fail(FCC_ERR_ARGUMENT, "unknown precision");
    if (!b->x || !b->y || !b->theta || !b->coef || !b->R0) return fail(FCC_ERR_DIMENSION, "missing body arrays");
    const int64_t N = cfg->n_bodies;
    for (int64_t i = 0; i < N; i++) {
        int32_t t = b->type ? b->type[i] : 1;
        if (t < 1 || t > cfg->n_types) return fail(FCC_ERR_ARGUMENT, "body type without BodyType entry");
    }
    for (int64_t i = 0; i < N; i++) {
        int32_t p = b->parent ? b->parent[i] : (int32_t)(i + 1);
        if (p < 1 || p > N)
            return fail(FCC_ERR_ARGUMENT, "body %lld has parent %d outside 1:%lld", (long long)(i + 1), p, (long long)N);
        int32_t t = b->type ? b->type[i] : 1;
        if (!cfg->types[t - 1].parent_drag) continue;
        if (p == i + 1)
            return fail(FCC_ERR_ARGUMENT, "body %lld has drag = :parent but is its own parent", (long long)(i + 1));
        int32_t tp = b->type ? b->type[p - 1] : 1;
        if (cfg->types[tp - 1].parent_drag)
            return fail(FCC_ERR_ARGUMENT, "body %lld has drag = :parent, but so has its parent %d (one level only)",
                        (long long)(i + 1), p);
    }
    int nbody = 0;
    for (int q = 0; q < cfg->n_terms; q++) {
        const int k = cfg->terms[q].kind;
        if (k == FCC_TERM_TENSION || k == FCC_TERM_PERIMETER_ELASTICITY || k == FCC_TERM_BENDING)
            nbody++;
        else if (k != FCC_TERM_OVERLAP && k != FCC_TERM_LINEAR_REPULSION && k != FCC_TERM_HARMONIC_BOND)
            return fail(FCC_ERR_ARGUMENT, "energy terms must be BodyTerms, ContactTerms or BondTerms");
    }
    if (nbody > MAXB) return fail(FCC_ERR_UNSUPPORTED, "at most %d native body terms", MAXB);
    return FCC_OK;
}

static uint32_t type_mask(const fcc_term &t) {
    if (t.n_on <= 0) return 0xFFFFFFFFu;
    uint32_t m = 0;
    for (int q = 0; q < t.n_on && q < FCC_MAX_TYPES; q++)
        if (t.on_types[q] >= 1 && t.on_types[q] <= 32) m |= 1u << (t.on_types[q] - 1);
    return m;
}

// ------------------------------------------------------- MFMA operand tables

static uint16_t to_f16_bits(double x) {
    const _Float16 h = (_Float16)(float)x;
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}
static double f16_value(uint16_t u) {
    _Float16 h;
    std::memcpy(&h, &u, 2);
    return (double)(float)h;
}
static uint16_t to_bf16_bits(double x) {  // round to nearest even
    const float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}
static double bf16_value(uint16_t u) {
    const uint32_t v = (uint32_t)u << 16;
    float f;
    std::memcpy(&f, &v, 4);
    return (double)f;
}

// Basis value of quantity q (0: C, 1: D, 2: DD) at node k for coefficient c of
// the reference order [a0, cos modes, sin modes] (src/basis.jl:46-58).
static double basis_value(int q, int k, int c, int K, int nth) {
    const double phi = (double)k * (2.0 * 3.141592653589793 / (double)nth);
    if (c == 0) return q == 0 ? 1.0 : 0.0;
    const int m = c <= K ? c : c - K;
    const int n = m + 1;
    const bool cosine = c <= K;
    const double cn = std::cos((double)n * phi), sn = std::sin((double)n * phi);
    if (q == 0) return cosine ? cn : sn;
    if (q == 1) return cosine ? (double)(-n) * sn : (double)n * cn;
    return cosine ? (double)(-n * n) * cn : (double)(-n * n) * sn;
}

// Physical meaning of row w (0..31) of output tile t of the MFMA layout:
// slot σ = 16t + r of lane half hh holds quantity σ / 20 at node 20·hh + σ % 20.
static bool row_phys(int t, int w, int T, int &q, int &k) {
    const int r = 4 * (w / 8) + (w % 4), hh = (w / 4) % 2;
    const int sig = 16 * t + r;
    const int nq = T == 4 ? 3 : 2;
    if (sig >= nq * NH) return false;
    q = sig / NH;
    k = NH * hh + sig % NH;
    return true;
}

// The same for row m (0..15) of output tile t of the four-lane layout: local
// slot s = 4t + m % 4 of lane group m / 4 holds quantity s / 10 at node
// 10·(m / 4) + s % 10.
static bool row_phys4(int t, int m, int T, int &q, int &k) {
    const int g = m / 4, s = 4 * t + m % 4;
    const int nq = T == 4 ? 3 : 2;
    if (s >= nq * NQ4) return false;
    q = s / NQ4;
    k = NQ4 * g + s % NQ4;
    return true;
}

// ------------------------------------------------------------- simulation

struct fcc_sim {
    bool borrowed = false;  // owned by an ensemble
    // The device status is the one the last fcc_run left, known on the host
    // (run() sets it; every entry point that may change the status or the
    // state clears it): the next run needs no start snapshot.
    bool mirror_ok = false;
    virtual ~fcc_sim() {}
    virtual int set_bodies(const fcc_bodies *b) = 0;
    virtual int get_bodies(fcc_bodies *b) = 0;
    virtual int set_dt(double dt) = 0;
    virtual int set_bonds(int64_t n, const int32_t *pairs) = 0;
    virtual int step(int64_t n) = 0;
    virtual int run(int64_t n) = 0;
    virtual int update_neighbors(int force) = 0;
    virtual int compute_gradients(int frame) = 0;
    virtual int get_gradients(double *, double *, double *, double *, double *, double *) = 0;
    virtual int get_contour_gradients(double *, double *, double *, double *) = 0;
    virtual int get_geometry(double *, double *, double *, double *, double *) = 0;
    virtual int shape_index(double *, double *, double *) = 0;
    virtual int get_celllist(int32_t *, int32_t *, int32_t *, int32_t *, int32_t *, int32_t *) = 0;
    virtual int get_neighbors(int32_t *, int32_t *, int32_t *) = 0;
    virtual int set_neighbor_capacity(int32_t) = 0;
    virtual int get_counters(int64_t *, double *, int64_t *, int64_t *) = 0;
    virtual int time_step_kernels(int64_t, double *, double *) = 0;
    virtual int run_timed(int64_t, double *, double *) = 0;
    virtual int time_kernels(int64_t, double *, double *) = 0;
    virtual int time_rebuild(int64_t, double *) = 0;
    virtual int contours(double *, double *) = 0;
    virtual int frame_request() = 0;
    virtual int frame_read(double *, double *, int64_t *) = 0;
    virtual int checkpoint(void *, int64_t, int64_t *) = 0;
    virtual int restore(const void *, int64_t) = 0;
    virtual int observer_start(const fcc_observer_config *) = 0;
    virtual int observe() = 0;
    virtual int observer_read(int32_t *, int64_t *, double *, double *, double *, uint64_t *) = 0;
};

template <typename T>
struct Sim final : fcc_sim {
    fcc_config cfg;
    Params<T> p{};
    Ptrs<T> d{};
    hipStream_t stream = nullptr;
    bool own_stream = true;
    int64_t N = 0;
    int K = 0, nc = 0, ldc = 0, nth = 0, KM = 0;
    bool mixed = false, debug = false, has_contact = false, has_pdrag = false, has_linrep = false;
    // Systems below this many bodies get the four-lane forces kernels at
    // nθ = 40 (FOUR_LANES_BELOW; 0 with FCC_FLAG_TWO_LANES).
    int64_t four_below = FOUR_LANES_BELOW;
    bool has_bonds = false;  // bond terms and at least one bond
    // One-kernel steps (k_step_mfma, k_step_fast2): a step writes positions,
    // lab rows and radius bounds into the other buffer of each pair, so d.x,
    // d.y, d.lab, d.rmax, d.rmin name the current state and d.xn, ... the
    // other buffers; `par` is the parity of the current buffers (flip()).
    bool member = false;       // replica of an ensemble (the ensemble launches its kernels)
    bool two_kernels = false;  // FCC_FLAG_TWO_KERNELS
    int par = 0;
    int64_t nbonds = 0;
    double skin = 0.0, range = 0.0;
    int32_t ncx = 1, ncy = 1;
    double wx = 0, wy = 0;
    int64_t nce = 1;
    int32_t *cellid = nullptr, *perm = nullptr, *cstart = nullptr, *ccnt = nullptr, *cfill = nullptr;
    int32_t *ctile = nullptr;  // per-tile totals of the cell-count scan
    double *xs = nullptr, *ys = nullptr;  // positions and bounds in cell-sorted order (list builds)
    T *rs = nullptr;
    unsigned ntile = 1;
    uint16_t *asyn = nullptr, *bpull = nullptr;  // MFMA operand tables (mixed modes)
    uint16_t *asyn4 = nullptr, *bpull4 = nullptr;  // the same for the four-lane layout
    int mfma_T = 0;
    bool valid = false;
    int32_t maxcount = 0;  // largest neighbour count of the last build
    int64_t builds = 0, nstep = 0;
    double t = 0.0;
    std::vector<void *> allocs;
    // graph mode (ext/FourierContourCellsCUDAExt.jl:52-71)
    // captured steps: gx[q][b] holds 2^b steps from the buffers of parity q
    // (b < NGRAPH: 1 to 32 steps; set 1 only for one-kernel steps, gone)
    static constexpr int NGRAPH = 6;
    hipGraphExec_t gx[2][NGRAPH] = {};
    bool gone = false;
    int gen = 0, gexec_gen = -1;
    bool want_timed = false;
    // Pipelined chunks of fcc_run: status snapshots in pinned memory, one
    // slot per queued chunk (and slot NQ for the status at the start of a
    // run), each with its completion event, and the timing events of each
    // chunk's eagerly launched last step (fcc_run_timed).
    static constexpr int NQ = 2;
    DevStatus *snap = nullptr, *snap_d = nullptr;  // pinned status snapshots, and their device address
    DevStatus mirror{};                            // the status a run left (mirror_ok)
    hipEvent_t sev[NQ + 1] = {};
    hipEvent_t tev[NQ][4] = {};
    double interval = 16.0;
    int64_t since_rebuild = 0, gbuilds = 0;

    ~Sim() override {
        if (stream) (void)hipStreamSynchronize(stream);
        if (fr.copy) (void)hipStreamSynchronize(fr.copy);
        if (fr.ready) (void)hipEventDestroy(fr.ready);
        if (fr.done) (void)hipEventDestroy(fr.done);
        if (fr.hX) (void)hipHostFree(fr.hX);
        if (fr.copy) (void)hipStreamDestroy(fr.copy);
        for (auto &set : gx)
            for (hipGraphExec_t g : set)
                if (g) (void)hipGraphExecDestroy(g);
        for (hipEvent_t e : sev)
            if (e) (void)hipEventDestroy(e);
        for (int q = 0; q < NQ; q++)
            for (hipEvent_t e : tev[q])
                if (e) (void)hipEventDestroy(e);
        if (snap) (void)hipHostFree(snap);
        for (void *q : allocs) (void)hipFree(q);
        if (stream && own_stream) (void)hipStreamDestroy(stream);
    }

    template <typename U>
    int alloc(U **out, size_t n) {
        void *q = nullptr;
        HIPCHK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(U)));
        allocs.push_back(q);
        *out = (U *)q;
        return FCC_OK;
    }
    template <typename U>
    void release(U *q) {
        if (!q) return;
        auto it = std::find(allocs.begin(), allocs.end(), (void *)q);
        if (it != allocs.end()) allocs.erase(it);
        (void)hipFree(q);
    }

    // `ext_stream` / `ext_st`: a stream and status slot owned by an ensemble.
    int init(const fcc_config *c, const fcc_bodies *b, hipStream_t ext_stream = nullptr,
             DevStatus *ext_st = nullptr) {
        cfg = *c;
        N = c->n_bodies;
        K = c->n_modes;
        nc = 2 * K + 1;
        ldc = nc <= 16 ? 16 : (nc <= 32 ? 32 : 64);  // row stride ≥ nc (2·31 + 1 at most)
        nth = c->n_theta;
        KM = K <= 5 ? 5 : (K <= 7 ? 7 : (K <= 15 ? 15 : 31));
        mixed = c->precision == FCC_MIXED_FP16 || c->precision == FCC_MIXED_BF16;
        debug = (c->flags & FCC_FLAG_DEBUG_BUFFERS) != 0;
        four_below = (c->flags & FCC_FLAG_TWO_LANES) ? 0 : FOUR_LANES_BELOW;
        two_kernels = (c->flags & FCC_FLAG_TWO_KERNELS) != 0;
        member = ext_st != nullptr;
        skin = c->skin;
        if (ext_stream) {
            stream = ext_stream;
            own_stream = false;
        } else {
            HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        }

        p.N = N;
        p.K = K;
        p.nc = nc;
        p.ldc = ldc;
        p.nth = nth;
        const double dphi = 2.0 * 3.141592653589793 / (double)nth;
        p.dphi = (T)dphi;
        p.Lx = c->Lx;
        p.Ly = c->Ly;
        p.invLx = 1.0 / c->Lx;
        p.invLy = 1.0 / c->Ly;
        // terms (src/simulation.jl:73-76)
        int nb = 0;
        int overlaps[FCC_MAX_TERMS];
        int nov = 0;
        bool linrep = false;
        for (int q = 0; q < c->n_terms; q++) {
            const fcc_term &tm = c->terms[q];
            if (tm.kind == FCC_TERM_OVERLAP) {
                overlaps[nov++] = q;
                continue;
            }
            if (tm.kind == FCC_TERM_HARMONIC_BOND) {
                p.bdmask[p.nbondt] = type_mask(tm);
                p.bdk[p.nbondt] = (T)tm.p[0];
                p.bdr0[p.nbondt] = (T)tm.p[1];
                p.nbondt++;
                continue;
            }
            if (tm.kind == FCC_TERM_LINEAR_REPULSION) {
                // g(d) = k·d for d > 0 (src/energies/contact.jl:39-41), summed per type of body i
                for (int t = 0; t < FCC_MAX_TYPES; t++)
                    if (type_mask(tm) & (1u << t)) p.krep[t] += (T)tm.p[0];
                linrep = true;
                continue;
            }
            p.bkind[nb] = tm.kind;
            p.bmask[nb] = type_mask(tm);
            p.bp0[nb] = (T)tm.p[0];
            p.bp1[nb] = (T)tm.p[1];
            if (tm.kind == FCC_TERM_BENDING) p.need_dd = 1;
            nb++;
        }
        p.nbody = nb;
        has_contact = nov > 0 || linrep;
        has_linrep = linrep;
        p.ncontact = has_contact ? 1 : 0;
        // combined Overlap table: right fold over the terms (contact.jl:46-51)
        for (int rel = 0; rel < 2; rel++)
            for (int ti = 0; ti < FCC_MAX_TYPES; ti++)
                for (int tj = 0; tj < FCC_MAX_TYPES; tj++) {
                    double acc = 0.0;
                    for (int q = nov - 1; q >= 0; q--) {
                        const fcc_term &tm = c->terms[overlaps[q]];
                        double g = 0.0;
                        if (type_mask(tm) & (1u << ti))
                            g = rel ? tm.Jrel[ti * FCC_MAX_TYPES + tj] : tm.J[ti * FCC_MAX_TYPES + tj];
                        acc = g + acc;
                    }
                    p.Jc[rel][ti][tj] = (T)acc;
                }
        // coupling classes of the fast forces kernel
        {
            std::vector<double> vals;
            for (int rel = 0; rel < 2; rel++)
                for (int ti = 0; ti < FCC_MAX_TYPES; ti++)
                    for (int tj = 0; tj < FCC_MAX_TYPES; tj++) {
                        const double v = (double)p.Jc[rel][ti][tj];
                        int cls = -1;
                        if (v != 0.0 && ti < c->n_types && tj < c->n_types) {
                            auto it = std::find(vals.begin(), vals.end(), v);
                            cls = (int)(it - vals.begin());
                            if (it == vals.end()) vals.push_back(v);
                        }
                        p.ccls[rel][ti][tj] = cls;
                    }
            p.ncls = (int)vals.size();
            for (size_t q = 0; q < vals.size(); q++) p.Jcls[q] = (T)vals[q];
            p.cls_uniform = 1;
            for (int rel = 0; rel < 2; rel++)
                for (int ti = 0; ti < c->n_types; ti++)
                    for (int tj = 0; tj < c->n_types; tj++)
                        if (p.ccls[rel][ti][tj] != 0) p.cls_uniform = 0;
        }
        range = 0.0;  // interaction_range(Overlap) = 0
        p.range = (T)range;
        for (int q = 0; q < FCC_MAX_TYPES; q++) {
            const fcc_body_type &bt = c->types[q];
            p.v0[q] = (T)bt.v0;
            p.Dr[q] = (T)bt.Dr;
            p.Dt[q] = (T)bt.Dt;
            p.mur[q] = (T)bt.mur;
            p.muth[q] = (T)bt.mutheta;
            p.muab[q] = (T)bt.muab;
            p.pdrag[q] = q < c->n_types ? bt.parent_drag : 0;
            if (q < c->n_types && bt.parent_drag) has_pdrag = true;
        }
        p.dt = (T)c->dt;
        p.one_type = c->n_types == 1 ? 1 : 0;
        noise_amplitudes();
        p.key0 = (uint32_t)c->seed;
        p.key1 = (uint32_t)(c->seed >> 32);
        p.skin = (T)skin;

        // buffers
        int rc = 0;
        if ((rc = alloc(&d.x, N)) || (rc = alloc(&d.y, N)) || (rc = alloc(&d.imgx, N)) || (rc = alloc(&d.imgy, N)) ||
            (rc = alloc(&d.th, N)) || (rc = alloc(&d.coef, N * ldc)) || (rc = alloc(&d.R0, N)) ||
            (rc = alloc(&d.type, N)) || (rc = alloc(&d.parent, N)) || (rc = alloc(&d.lab, N * ldc)) ||
            (rc = alloc(&d.rmax, N)) || (rc = alloc(&d.rmin, N)) || (rc = alloc(&d.glab, N * ldc)) || (rc = alloc(&d.gx, N)) ||
            (rc = alloc(&d.gy, N)) || (rc = alloc(&d.gth, N)) || (rc = alloc(&d.Ebody, N)) ||
            (rc = alloc(&d.gcoef, N * ldc)))
            return rc;
        if (ext_st)
            d.st = ext_st;
        else if ((rc = alloc(&d.st, 1)))
            return rc;
        HIPCHK(hipMemsetAsync(d.coef, 0, sizeof(T) * N * ldc, stream));
        HIPCHK(hipMemsetAsync(d.lab, 0, sizeof(T) * N * ldc, stream));
        HIPCHK(hipMemsetAsync(d.glab, 0, sizeof(T) * N * ldc, stream));
        HIPCHK(hipMemsetAsync(d.gcoef, 0, sizeof(T) * N * ldc, stream));
        HIPCHK(hipMemsetAsync(d.gx, 0, sizeof(T) * N, stream));
        HIPCHK(hipMemsetAsync(d.gy, 0, sizeof(T) * N, stream));
        HIPCHK(hipMemsetAsync(d.gth, 0, sizeof(T) * N, stream));
        HIPCHK(hipMemsetAsync(d.Ebody, 0, sizeof(T) * N, stream));
        HIPCHK(hipMemsetAsync(d.st, 0, sizeof(DevStatus), stream));
        hipLaunchKernelGGL(k_clear_status, 1, 1, 0, stream, d.st);
        hipLaunchKernelGGL(k_set_target, 1, 1, 0, stream, d.st, INT64_MAX);
        if (debug) {
            if ((rc = alloc(&d.P, N * nth)) || (rc = alloc(&d.GR, N * nth)) || (rc = alloc(&d.GD, N * nth)) ||
                (rc = alloc(&d.GDD, N * nth)))
                return rc;
            HIPCHK(hipMemsetAsync(d.P, 0, sizeof(T) * N * nth, stream));
        }
        // basis (src/basis.jl:35-63) in the device row layout
        std::vector<T> S((size_t)3 * nth * SB, T(0)), cs((size_t)2 * nth);
        for (int k = 0; k < nth; k++) {
            const double phi = (double)k * dphi;
            cs[2 * k] = (T)std::cos(phi);
            cs[2 * k + 1] = (T)std::sin(phi);
            S[(size_t)k * SB] = T(1);
            for (int m = 1; m <= K; m++) {
                const int n = m + 1;
                const double sn = std::sin((double)n * phi), cn = std::cos((double)n * phi);
                S[(size_t)k * SB + m] = (T)cn;
                S[(size_t)k * SB + SH + m] = (T)sn;
                S[(size_t)(nth + k) * SB + m] = (T)((double)(-n) * sn);
                S[(size_t)(nth + k) * SB + SH + m] = (T)((double)n * cn);
                S[(size_t)(2 * nth + k) * SB + m] = (T)((double)(-n * n) * cn);
                S[(size_t)(2 * nth + k) * SB + SH + m] = (T)((double)(-n * n) * sn);
            }
        }
        std::vector<T> Sc((size_t)3 * nth * ldc, T(0));
        for (int q = 0; q < 3; q++)
            for (int k = 0; k < nth; k++) {
                const T *row = &S[(size_t)(q * nth + k) * SB];
                T *out = &Sc[(size_t)(q * nth + k) * ldc];
                out[0] = row[0];
                for (int m = 1; m <= K; m++) {
                    out[m] = row[m];
                    out[K + m] = row[SH + m];
                }
            }
        T *dS, *dcs, *dSc;
        if ((rc = alloc(&dS, S.size())) || (rc = alloc(&dcs, cs.size())) || (rc = alloc(&dSc, Sc.size()))) return rc;
        HIPCHK(hipMemcpyAsync(dSc, Sc.data(), sizeof(T) * Sc.size(), hipMemcpyHostToDevice, stream));
        d.Sc = dSc;
        HIPCHK(hipMemcpyAsync(dS, S.data(), sizeof(T) * S.size(), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(dcs, cs.data(), sizeof(T) * cs.size(), hipMemcpyHostToDevice, stream));
        d.S = dS;
        d.cs = dcs;
        if (mixed && nth == NTH_M && (K == 5 || K == 7)) {
            const bool bf = c->precision == FCC_MIXED_BF16;
            const int TT = p.need_dd ? 4 : 3;
            auto enc = [&](double x, uint16_t &hi, uint16_t &lo) {
                hi = bf ? to_bf16_bits(x) : to_f16_bits(x);
                const double rem = x - (bf ? bf16_value(hi) : f16_value(hi));
                lo = bf ? to_bf16_bits(rem) : to_f16_bits(rem);
            };
            std::vector<uint16_t> A((size_t)TT * 2 * 64 * 8, 0), B((size_t)2 * TT * 2 * 64 * 8, 0);
            for (int t = 0; t < TT; t++)
                for (int l = 0; l < 64; l++)
                    for (int sl = 0; sl < 8; sl++) {
                        int q, k;
                        const int cc = 8 * (l / 32) + sl;
                        double v = 0.0;
                        if (cc < nc && row_phys(t, l % 32, TT, q, k)) v = basis_value(q, k, cc, K, nth);
                        uint16_t hi, lo;
                        enc(v, hi, lo);
                        A[((size_t)(t * 2 + 0) * 64 + l) * 8 + sl] = hi;
                        A[((size_t)(t * 2 + 1) * 64 + l) * 8 + sl] = lo;
                    }
            for (int ch = 0; ch < 2 * TT; ch++)
                for (int l = 0; l < 64; l++)
                    for (int sl = 0; sl < 8; sl++) {
                        const int n = l % 32, hh = l / 32, t = ch / 2;
                        const int w = 16 * (ch % 2) + 8 * (sl / 4) + 4 * hh + (sl % 4);
                        int q, k;
                        const int coefi = n < 16 ? n : n - 16;
                        double v = 0.0;
                        if (coefi < nc && row_phys(t, w, TT, q, k)) v = basis_value(q, k, coefi, K, nth);
                        uint16_t hi, lo;
                        enc(v, hi, lo);
                        B[((size_t)(ch * 2 + 0) * 64 + l) * 8 + sl] = n < 16 ? hi : lo;
                        B[((size_t)(ch * 2 + 1) * 64 + l) * 8 + sl] = n < 16 ? hi : 0;
                    }
            if ((rc = alloc(&asyn, A.size())) || (rc = alloc(&bpull, B.size()))) return rc;
            HIPCHK(hipMemcpy(asyn, A.data(), A.size() * 2, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(bpull, B.data(), B.size() * 2, hipMemcpyHostToDevice));
            // four-lane layout (16x16x16 tiles): synthesis A holds row l % 16 of
            // tile t at coefficients 4(l / 16) + j; pullback B holds row
            // 4(l / 16) + j of tile t at coefficient l % 16
            const int T4 = TT == 4 ? 8 : 5;
            std::vector<uint16_t> A4((size_t)T4 * 2 * 64 * 4, 0), B4((size_t)T4 * 2 * 64 * 4, 0);
            for (int t = 0; t < T4; t++)
                for (int l = 0; l < 64; l++)
                    for (int j = 0; j < 4; j++) {
                        int q, k;
                        uint16_t hi, lo;
                        const int ca = 4 * (l / 16) + j;
                        double v = 0.0;
                        if (ca < nc && row_phys4(t, l % 16, TT, q, k)) v = basis_value(q, k, ca, K, nth);
                        enc(v, hi, lo);
                        A4[((size_t)(t * 2 + 0) * 64 + l) * 4 + j] = hi;
                        A4[((size_t)(t * 2 + 1) * 64 + l) * 4 + j] = lo;
                        const int cb = l % 16;
                        v = 0.0;
                        if (cb < nc && row_phys4(t, 4 * (l / 16) + j, TT, q, k)) v = basis_value(q, k, cb, K, nth);
                        enc(v, hi, lo);
                        B4[((size_t)(t * 2 + 0) * 64 + l) * 4 + j] = hi;
                        B4[((size_t)(t * 2 + 1) * 64 + l) * 4 + j] = lo;
                    }
            if ((rc = alloc(&asyn4, A4.size())) || (rc = alloc(&bpull4, B4.size()))) return rc;
            HIPCHK(hipMemcpy(asyn4, A4.data(), A4.size() * 2, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(bpull4, B4.data(), B4.size() * 2, hipMemcpyHostToDevice));
            if (ldc != LDC_M) return fail(FCC_ERR_STATE, "MFMA forces kernel expects coefficient rows of %d", LDC_M);
            mfma_T = TT;
        }

        // cell grid (src/simulation.jl:90, src/neighbors.jl:27-33)
        double maxR0 = 0.0;
        for (int64_t i = 0; i < N; i++) maxR0 = std::max(maxR0, b->R0[i]);
        const double cutoff = 2.0 * maxR0 * (1.0 + std::sqrt(2.0 / 3.0)) + range + skin;
        auto ncells = [&](double L) {
            int n = (int)std::floor(L / cutoff);
            return n >= 3 ? n : 1;
        };
        ncx = ncells(c->Lx);
        ncy = ncells(c->Ly);
        wx = c->Lx / ncx;
        wy = c->Ly / ncy;
        nce = (int64_t)ncx * ncy;
        // node offsets |dxij + R c| stay below L/2 when L > 4 cutoff
        p.small_box = (c->Lx > 4.0 * cutoff && c->Ly > 4.0 * cutoff) ? 0 : 1;
        ntile = blocks(nce, SCAN_TILE);
        if ((rc = alloc(&cellid, N)) || (rc = alloc(&perm, N)) || (rc = alloc(&cstart, nce)) ||
            (rc = alloc(&ccnt, nce)) || (rc = alloc(&cfill, nce)) || (rc = alloc(&ctile, ntile)) ||
            (rc = alloc(&xs, N)) || (rc = alloc(&ys, N)) || (rc = alloc(&rs, N)))
            return rc;
        HIPCHK(hipMemsetAsync(cellid, 0, sizeof(int32_t) * N, stream));
        HIPCHK(hipMemsetAsync(perm, 0, sizeof(int32_t) * N, stream));
        HIPCHK(hipMemsetAsync(cstart, 0, sizeof(int32_t) * nce, stream));
        HIPCHK(hipMemsetAsync(ccnt, 0, sizeof(int32_t) * nce, stream));
        // neighbour list (capacity 32, src/neighbors.jl:121-127)
        d.cap = 32;
        if ((rc = alloc(&d.nidx, (size_t)d.cap * N)) || (rc = alloc(&d.ncount, N)) || (rc = alloc(&d.x0, N)) ||
            (rc = alloc(&d.y0, N)) || (rc = alloc(&d.rmax0, N)))
            return rc;
        HIPCHK(hipMemsetAsync(d.nidx, 0, sizeof(int32_t) * d.cap * N, stream));
        HIPCHK(hipMemsetAsync(d.ncount, 0, sizeof(int32_t) * N, stream));
        // children (src/bodies.jl:121-135), slot-major
        std::vector<int32_t> cnt(N, 0);
        for (int64_t k = 0; k < N; k++) {
            int32_t pr = (b->parent ? b->parent[k] : (int32_t)(k + 1)) - 1;
            if (pr != k) cnt[pr]++;
        }
        int32_t cmax = 1;
        for (int64_t k = 0; k < N; k++) cmax = std::max(cmax, cnt[k]);
        d.ccap = cmax;
        std::vector<int32_t> cidx((size_t)cmax * N, 0);
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int64_t k = 0; k < N; k++) {
            int32_t pr = (b->parent ? b->parent[k] : (int32_t)(k + 1)) - 1;
            if (pr == k) continue;
            cidx[(size_t)cnt[pr] * N + pr] = (int32_t)k;
            cnt[pr]++;
        }
        if ((rc = alloc(&d.cidx, cidx.size())) || (rc = alloc(&d.ccount, N))) return rc;
        HIPCHK(hipMemcpy(d.cidx, cidx.data(), sizeof(int32_t) * cidx.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d.ccount, cnt.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice));
        if (has_pdrag && (rc = alloc(&d.th0, N))) return rc;
        return set_bodies(b);
    }

    int launch_init() {
        const unsigned g = blocks(N, 256);
        switch (KM) {
        case 5:
            if (mixed) hipLaunchKernelGGL((k_init<T, 5, true>), g, 256, 0, stream, p, d);
            else hipLaunchKernelGGL((k_init<T, 5, false>), g, 256, 0, stream, p, d);
            break;
        case 7:
            if (mixed) hipLaunchKernelGGL((k_init<T, 7, true>), g, 256, 0, stream, p, d);
            else hipLaunchKernelGGL((k_init<T, 7, false>), g, 256, 0, stream, p, d);
            break;
        case 15:
            if (mixed) hipLaunchKernelGGL((k_init<T, 15, true>), g, 256, 0, stream, p, d);
            else hipLaunchKernelGGL((k_init<T, 15, false>), g, 256, 0, stream, p, d);
            break;
        default:
            if (mixed) hipLaunchKernelGGL((k_init<T, 31, true>), g, 256, 0, stream, p, d);
            else hipLaunchKernelGGL((k_init<T, 31, false>), g, 256, 0, stream, p, d);
        }
        HIPCHK(hipGetLastError());
        return FCC_OK;
    }

    // The fast kernel needs K in {5, 7}, at most 64 nodes, at most two coupling
    // classes and neighbour lists shorter than 2^NPLANE (its per-node hit
    // counters).
    bool fast_ok() const {
        return (K == 5 || K == 7) && nth <= FUSED_MAX_THETA && p.ncls <= 2 && !has_linrep &&
               (!has_contact || maxcount < (1 << NPLANE));
    }

    // Forces of one step: the forces kernel, then the bond kernel when bond
    // terms act on bonds (compute_gradients!, src/simulation.jl:150-158).
    int launch_forces(int use_status, hipEvent_t eb = nullptr, hipEvent_t ee = nullptr) {
        int rc = launch_forces_kernel(use_status, eb, ee);
        if (rc || !has_bonds) return rc;
        hipLaunchKernelGGL((k_bonds<T>), blocks(N, 256), 256, 0, stream, p, d, use_status ? 1 : 0);
        HIPCHK(hipGetLastError());
        return FCC_OK;
    }

    int launch_forces_kernel(int use_status, hipEvent_t eb, hipEvent_t ee) {
        const unsigned g = blocks(N, 64);
        const int dbg = debug ? 1 : 0;
        if constexpr (sizeof(T) == 4) {
            if (mfma_T && fast_ok()) {
                const unsigned gm = blocks(N, 32);
                const bool bf = cfg.precision == FCC_MIXED_BF16;
// Below FOUR_LANES_BELOW bodies the four-lane kernel (16 bodies per wave).
#define FCC_MF2(KV, TV, NCV, BFV)                                                                        \
    if (N < four_below)                                                                            \
        FCC_LAUNCH((k_forces_mfma4<KV, TV, NCV, BFV>), blocks(N, 16), 64, 0, stream, eb, ee, p, d, asyn4, bpull4, use_status, dbg); \
    else FCC_LAUNCH((k_forces_mfma<KV, TV, NCV, BFV>), gm, 64, 0, stream, eb, ee, p, d, asyn, bpull, use_status, dbg);
#define FCC_MF(KV, TV, NCV)                                                                             \
    if (bf) { FCC_MF2(KV, TV, NCV, true) } else { FCC_MF2(KV, TV, NCV, false) }
#define FCC_MF_T(KV, NCV) if (mfma_T == 3) { FCC_MF(KV, 3, NCV) } else { FCC_MF(KV, 4, NCV) }
                if (K == 5) {
                    if (p.ncls <= 1) { FCC_MF_T(5, 1) } else { FCC_MF_T(5, 2) }
                } else {
                    if (p.ncls <= 1) { FCC_MF_T(7, 1) } else { FCC_MF_T(7, 2) }
                }
#undef FCC_MF_T
#undef FCC_MF
#undef FCC_MF2
                HIPCHK(hipGetLastError());
                return FCC_OK;
            }
        }
        if (fast_ok()) {
// Below FOUR_LANES_BELOW bodies (fewer than 4096 waves of 32 bodies, one per
// SIMD slot of occupancy 4) the FP32 / FP64 kernel takes four lanes per body.
#define FCC_FAST(KV, NC)                                                                                 \
    if (nth == 40 && N < four_below)                                                               \
        FCC_LAUNCH((k_forces_fast2<T, KV, NC, 4>), blocks(N, 16), 64, 0, stream, eb, ee, p, d, use_status, dbg); \
    else if (nth == 40)                                                                                  \
        FCC_LAUNCH((k_forces_fast2<T, KV, NC, 2>), blocks(N, 32), 64, 0, stream, eb, ee, p, d, use_status, dbg); \
    else FCC_LAUNCH((k_forces_fast<T, 0, KV, NC>), g, 64, 0, stream, eb, ee, p, d, use_status, dbg);
            if (K == 5) {
                if (p.ncls <= 1) { FCC_FAST(5, 1) } else { FCC_FAST(5, 2) }
            } else {
                if (p.ncls <= 1) { FCC_FAST(7, 1) } else { FCC_FAST(7, 2) }
            }
#undef FCC_FAST
            HIPCHK(hipGetLastError());
            return FCC_OK;
        }
        if (nth > FUSED_MAX_THETA) {
            const size_t sh = sizeof(T) * (size_t)nth;  // contact density of the body's nodes
#define FCC_WIDE(KMV)                                                                                    \
    if (nth <= WIDE_WG64_THETA)                                                                          \
        FCC_LAUNCH((k_forces_wide<T, KMV, 64>), (unsigned)N, 64, sh, stream, eb, ee, p, d, use_status, dbg); \
    else FCC_LAUNCH((k_forces_wide<T, KMV, 256>), (unsigned)N, 256, sh, stream, eb, ee, p, d, use_status, dbg);
            switch (KM) {
            case 5: FCC_WIDE(5) break;
            case 7: FCC_WIDE(7) break;
            case 15: FCC_WIDE(15) break;
            default: FCC_WIDE(31)
            }
#undef FCC_WIDE
            HIPCHK(hipGetLastError());
            return FCC_OK;
        }
#define FCC_FORCES(KMV)                                                                                  \
    if (nth == 40) FCC_LAUNCH((k_forces<T, 40, KMV>), g, 64, 0, stream, eb, ee, p, d, use_status, dbg);   \
    else FCC_LAUNCH((k_forces<T, 0, KMV>), g, 64, 0, stream, eb, ee, p, d, use_status, dbg);
        switch (KM) {
        case 5: FCC_FORCES(5) break;
        case 7: FCC_FORCES(7) break;
        case 15: FCC_FORCES(15) break;
        default: FCC_FORCES(31)
        }
#undef FCC_FORCES
        HIPCHK(hipGetLastError());
        return FCC_OK;
    }

    int launch_update(hipEvent_t eb = nullptr, hipEvent_t ee = nullptr) {
        const unsigned g = blocks(N, 256);
        const int pd = has_pdrag ? 1 : 0, dbg = debug ? 1 : 0;
#define FCC_UPD(KMV, EX)                                                                                   \
    if (mixed) FCC_LAUNCH((k_update<T, KMV, EX, true>), g, 256, 0, stream, eb, ee, p, d, pd, dbg);          \
    else FCC_LAUNCH((k_update<T, KMV, EX, false>), g, 256, 0, stream, eb, ee, p, d, pd, dbg);
        if (K == 5) {
            FCC_UPD(5, true)
        } else if (K == 7) {
            FCC_UPD(7, true)
        } else if (KM == 5) {
            FCC_UPD(5, false)
        } else if (KM == 7) {
            FCC_UPD(7, false)
        } else if (KM == 15) {
            FCC_UPD(15, false)
        } else {
            FCC_UPD(31, false)
        }
#undef FCC_UPD
        HIPCHK(hipGetLastError());
        return FCC_OK;
    }

    // Whether a step runs as one kernel: single simulations on the two- or
    // four-lane FP32 / mixed kernels (the mixed ones without Bending) with
    // contact and a Verlet skin, without parent drag (a child reads its
    // parent's θ before the update), bonds (their kernel adds to the centre
    // forces between the forces and the update) or debug buffers.
    bool one_kernel() const {
        if constexpr (sizeof(T) != 4) return false;
        return !member && !two_kernels && fast_ok() && nth == 40 && has_contact && skin > 0.0 && !has_pdrag &&
               !has_bonds && !debug && (!mfma_T || mfma_T == 3);
    }

    // The other buffer of each state pair, allocated at the first one-kernel step.
    int ensure_next() {
        if (d.xn) return FCC_OK;
        int rc;
        if ((rc = alloc(&d.xn, N)) || (rc = alloc(&d.yn, N)) || (rc = alloc(&d.labn, N * ldc)) ||
            (rc = alloc(&d.rmaxn, N)) || (rc = alloc(&d.rminn, N)))
            return rc;
        HIPCHK(hipMemsetAsync(d.labn, 0, sizeof(T) * N * ldc, stream));
        return FCC_OK;
    }

    // Before one-kernel steps from the current buffers (and after eager ones):
    // a skin trip of the last one-kernel step becomes the rebuild flag, and
    // the step clock slot of parity `par` takes the clock (k_step_slots).
    int sync_slots(DevStatus *out = nullptr) {
        hipLaunchKernelGGL(k_step_slots, 1, 1, 0, stream, d.st, par, out);
        HIPCHK(hipGetLastError());
        return FCC_OK;
    }

    // The state of an executed one-kernel step is in the other buffers.
    void flip() {
        std::swap(d.x, d.xn);
        std::swap(d.y, d.yn);
        std::swap(d.lab, d.labn);
        std::swap(d.rmax, d.rmaxn);
        std::swap(d.rmin, d.rminn);
        par ^= 1;
    }
    void set_parity(int q) {
        if (par != q) flip();
    }

    // use_status 3: the step's status protocol (step_open, reading the
    // buffers of parity `par`); 4: the same without the freeze check (timing
    // series, whose state is restored afterwards).
    // Below four_below bodies the four-lane kernels (four waves of 16 bodies
    // per workgroup), else the two-lane ones (two waves of 32).
    int launch_one_kernel(int use_status, hipEvent_t eb, hipEvent_t ee) {
        const unsigned g = blocks(N, 64);
        const bool four = N < four_below;
        if constexpr (sizeof(T) == 4) {
            if (mfma_T) {
                const bool bf = cfg.precision == FCC_MIXED_BF16;
#define FCC_STEP_MF2(KV, NCV, BFV)                                                                        \
    if (four) FCC_LAUNCH((k_step_mfma4<KV, 3, NCV, BFV>), g, 256, 0, stream, eb, ee, p, d, asyn4, bpull4, use_status, par); \
    else FCC_LAUNCH((k_step_mfma<KV, 3, NCV, BFV>), g, 128, 0, stream, eb, ee, p, d, asyn, bpull, use_status, par);
#define FCC_STEP_MF(KV, NCV) if (bf) { FCC_STEP_MF2(KV, NCV, true) } else { FCC_STEP_MF2(KV, NCV, false) }
                if (K == 5) {
                    if (p.ncls <= 1) { FCC_STEP_MF(5, 1) } else { FCC_STEP_MF(5, 2) }
                } else {
                    if (p.ncls <= 1) { FCC_STEP_MF(7, 1) } else { FCC_STEP_MF(7, 2) }
                }
#undef FCC_STEP_MF
#undef FCC_STEP_MF2
            } else {
#define FCC_STEP_F(KV, NCV)                                                                               \
    if (four) FCC_LAUNCH((k_step_fast4<T, KV, NCV>), g, 256, 0, stream, eb, ee, p, d, use_status, par);     \
    else FCC_LAUNCH((k_step_fast2<T, KV, NCV>), g, 128, 0, stream, eb, ee, p, d, use_status, par);
                if (K == 5) {
                    if (p.ncls <= 1) { FCC_STEP_F(5, 1) } else { FCC_STEP_F(5, 2) }
                } else {
                    if (p.ncls <= 1) { FCC_STEP_F(7, 1) } else { FCC_STEP_F(7, 2) }
                }
#undef FCC_STEP_F
            }
            HIPCHK(hipGetLastError());
            return FCC_OK;
        }
        return fail(FCC_ERR_STATE, "one-kernel step in double precision");
    }

    // One step on the stream: one kernel where one_kernel() holds (the host
    // then flips to the buffers it writes), else the forces and the update
    // kernels. eb/ee time the forces kernel (the step's kernel in a
    // one-kernel step), ub/ue the update kernel (recorded back to back in a
    // one-kernel step). The other buffers must exist (ensure_next) before a
    // one-kernel step is captured.
    int launch_step(hipEvent_t eb = nullptr, hipEvent_t ee = nullptr, hipEvent_t ub = nullptr,
                    hipEvent_t ue = nullptr) {
        int rc;
        if (one_kernel()) {
            if ((rc = launch_one_kernel(3, eb, ee))) return rc;
            if (ub) HIPCHK(hipEventRecord(ub, stream));
            if (ue) HIPCHK(hipEventRecord(ue, stream));
            flip();
            return FCC_OK;
        }
        if ((rc = launch_forces(1, eb, ee))) return rc;
        return launch_update(ub, ue);
    }

    // advance! (src/simulation.jl:181-191) as device work on the stream. An
    // eager one-kernel step leaves a skin trip in the rebuild flag.
    int device_step() {
        p.skin_check = (has_contact && valid && skin > 0.0) ? 1 : 0;
        if (!one_kernel()) return launch_step();
        int rc;
        if ((rc = ensure_next()) || (rc = sync_slots()) || (rc = launch_step())) return rc;
        return sync_slots();
    }

    // update_celllist! + update_neighbors!(nl) (src/neighbors.jl:81-92,223-242).
    // `ev` (bench timing): six events recorded around the stages cell index,
    // scan, scatter, cell sort and neighbours of the first pass.
    // The stages of a timed build (ev non-null) record ev[2s] at the start of
    // stage s's first kernel and ev[2s + 1] at the end of its last one, by
    // the dispatches themselves (hipExtLaunchKernelGGL).
    int rebuild(hipEvent_t *ev = nullptr) {
        const unsigned gN = blocks(N, 256), gC = blocks(nce, 256);
        auto E = [&](int q) -> hipEvent_t { return ev ? ev[q] : nullptr; };
        FCC_LAUNCH(k_fill_i32, gC, 256, 0, stream, E(0), nullptr, nce, ccnt, 0);
        FCC_LAUNCH(k_cell_index, gN, 256, 0, stream, nullptr, E(1), N, d.x, d.y, ncx, ncy, wx, wy, cellid, ccnt);
        FCC_LAUNCH(k_scan_tiles, ntile, SCAN_TILE, 0, stream, E(2), nullptr, nce, ccnt, cstart, ctile);
        FCC_LAUNCH(k_scan_offsets, ntile, SCAN_TILE, 0, stream, nullptr, E(3), nce, ctile, cstart, cfill);
        FCC_LAUNCH(k_scatter, gN, 256, 0, stream, E(4), E(5), N, cellid, cfill, perm);
        FCC_LAUNCH((k_cell_sort<T>), gC, 256, 0, stream, E(6), E(7), nce, cstart, ccnt, perm, d.x, d.y, d.rmax, xs,
                   ys, rs);
        HIPCHK(hipGetLastError());
        double *x0 = skin > 0.0 ? d.x0 : nullptr, *y0 = skin > 0.0 ? d.y0 : nullptr;
        T *r0 = skin > 0.0 ? d.rmax0 : nullptr;
        for (bool first = true;; first = false) {
            hipLaunchKernelGGL(k_clear_status, 1, 1, 0, stream, d.st);
            FCC_LAUNCH((k_neighbors<T>), gN, 256, 0, stream, first ? E(8) : nullptr, first ? E(9) : nullptr, N, d.x,
                       d.y, d.rmax, perm, cstart, ccnt, xs, ys, rs, p.Lx, p.Ly, ncx, ncy, wx, wy, range + skin, d.cap,
                       d.nidx, d.ncount, d.st, x0, y0, r0);
            HIPCHK(hipGetLastError());
            DevStatus st;
            HIPCHK(hipMemcpyAsync(&st, d.st, sizeof st, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            if (st.overflow <= d.cap) {
                const bool fast_before = fast_ok();
                maxcount = st.overflow;
                if (fast_ok() != fast_before) gen++;  // the forces kernel baked into the graph changes
                break;
            }
            const int32_t cap = 2 * st.overflow;
            int32_t *ni = nullptr;
            int rc = alloc(&ni, (size_t)cap * N);
            if (rc) return rc;
            HIPCHK(hipMemsetAsync(ni, 0, sizeof(int32_t) * cap * N, stream));
            release(d.nidx);
            d.nidx = ni;
            d.cap = cap;
            gen++;
        }
        hipLaunchKernelGGL(k_clear_status, 1, 1, 0, stream, d.st);
        HIPCHK(hipGetLastError());
        valid = true;
        builds++;
        return FCC_OK;
    }

    // The same build enqueued on a side stream `sb` without waiting for it
    // (ensemble runs): the replica stays frozen by its rebuild flag, which
    // k_finish_build releases if the list fits its capacity and `limit`.
    int rebuild_async(hipStream_t sb, int32_t limit) {
        const unsigned gN = blocks(N, 256), gC = blocks(nce, 256);
        hipLaunchKernelGGL(k_fill_i32, gC, 256, 0, sb, nce, ccnt, 0);
        hipLaunchKernelGGL(k_cell_index, gN, 256, 0, sb, N, d.x, d.y, ncx, ncy, wx, wy, cellid, ccnt);
        hipLaunchKernelGGL(k_scan_tiles, ntile, SCAN_TILE, 0, sb, nce, ccnt, cstart, ctile);
        hipLaunchKernelGGL(k_scan_offsets, ntile, SCAN_TILE, 0, sb, nce, ctile, cstart, cfill);
        hipLaunchKernelGGL(k_scatter, gN, 256, 0, sb, N, cellid, cfill, perm);
        hipLaunchKernelGGL((k_cell_sort<T>), gC, 256, 0, sb, nce, cstart, ccnt, perm, d.x, d.y, d.rmax, xs, ys, rs);
        hipLaunchKernelGGL(k_begin_build, 1, 1, 0, sb, d.st);
        double *x0 = skin > 0.0 ? d.x0 : nullptr, *y0 = skin > 0.0 ? d.y0 : nullptr;
        T *r0 = skin > 0.0 ? d.rmax0 : nullptr;
        hipLaunchKernelGGL((k_neighbors<T>), gN, 256, 0, sb, N, d.x, d.y, d.rmax, perm, cstart, ccnt, xs, ys, rs,
                           p.Lx, p.Ly, ncx, ncy, wx, wy, range + skin, d.cap, d.nidx, d.ncount, d.st, x0, y0, r0);
        hipLaunchKernelGGL(k_finish_build, 1, 1, 0, sb, d.st, d.cap, limit);
        HIPCHK(hipGetLastError());
        return FCC_OK;
    }

    // update_neighbors!(sim; force) (src/simulation.jl:132-142). The skin
    // check of the current state is the flag raised by the fused check of the
    // last update (the same test on the same state).
    int update_neighbors(int force) override {
        if (!force && valid && skin > 0.0) {
            DevStatus st;
            HIPCHK(hipMemcpyAsync(&st, d.st, sizeof st, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            if (st.rebuild == 0) return FCC_OK;
        }
        return rebuild();
    }

    int step(int64_t n) override {
        for (int64_t q = 0; q < n; q++) {
            int rc;
            if (has_contact && (rc = update_neighbors(0))) return rc;
            if ((rc = device_step())) return rc;
            t += cfg.dt;
            nstep++;
        }
        HIPCHK(hipStreamSynchronize(stream));
        return FCC_OK;
    }

    // run_steps! in graph mode (ext/FourierContourCellsCUDAExt.jl:65-71,133-180).
    // Graphs of 1, 2, 4, ..., 32 steps in a row (a graph launch costs
    // several µs between replays; inside a graph the kernels follow each
    // other without a gap), so k steps take k / 32 + popcount(k mod 32)
    // launches.
    // A captured one-kernel step flips the host's buffers as an executed one
    // would; the parity the capture started from is restored.
    int capture_steps(int steps, hipGraphExec_t *out) {
        hipGraph_t graph;
        const int par0 = par;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = FCC_OK;
        for (int q = 0; q < steps && !rc; q++) rc = launch_step();
        hipError_t e = hipStreamEndCapture(stream, &graph);
        set_parity(par0);
        if (rc) return rc;
        HIPCHK(e);
        HIPCHK(hipGraphInstantiate(out, graph, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(graph));
        HIPCHK(hipGraphUpload(*out, stream));  // its first replay in a run starts at once
        return FCC_OK;
    }

    int capture() {
        std::lock_guard<std::mutex> lk(g_capture_lock);
        for (auto &set : gx)
            for (hipGraphExec_t &g : set)
                if (g) {
                    HIPCHK(hipGraphExecDestroy(g));
                    g = nullptr;
                }
        p.skin_check = (has_contact && skin > 0.0) ? 1 : 0;
        gone = one_kernel();
        if (gone) {
            int rc = ensure_next();
            if (rc) return rc;
        }
        // one-kernel steps: a set of graphs for each parity of the buffers
        const int par0 = par;
        for (int q = 0; q < (gone ? 2 : 1); q++) {
            if (gone) set_parity(q);
            for (int b = 0; b < NGRAPH; b++)
                if (int rc = capture_steps(1 << b, &gx[q][b])) {
                    set_parity(par0);
                    return rc;
                }
        }
        set_parity(par0);
        gexec_gen = gen;
        if (gone) return sync_slots();  // the steps before may have been two-kernel ones
        return FCC_OK;
    }

    // k graph steps: k / 32 launches of the longest graph, then one launch
    // per set bit of the rest.
    // With one-kernel steps the host flips its buffers as if every step
    // executes (an odd graph, the one-step one, flips them).
    int replay(int64_t k) {
        constexpr int64_t LONG = 1 << (NGRAPH - 1);
        for (int64_t q = 0; q < k / LONG; q++) HIPCHK(hipGraphLaunch(gx[gone ? par : 0][NGRAPH - 1], stream));
        for (int b = NGRAPH - 2; b >= 0; b--)
            if ((k % LONG) & (1 << b)) {
                HIPCHK(hipGraphLaunch(gx[gone ? par : 0][b], stream));
                if (gone && b == 0) flip();
            }
        return FCC_OK;
    }

    double acc_f = 0.0, acc_u = 0.0;
    int64_t acc_n = 0;

    // `reps` forced list rebuilds of the current state, timed per stage.
    int time_rebuild(int64_t reps, double *stage_ms) override {
        hipEvent_t ev[10] = {};
        for (auto &e : ev) HIPCHK(hipEventCreate(&e));
        double acc[5] = {0, 0, 0, 0, 0};
        int rc = FCC_OK;
        for (int64_t q = 0; q < reps && !rc; q++) {
            if ((rc = rebuild(ev))) break;
            for (int k = 0; k < 5; k++) {
                float ms = 0.f;
                hipError_t e = hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]);
                if (e != hipSuccess) {
                    rc = fail(FCC_ERR_DEVICE, "hipEventElapsedTime: %s", hipGetErrorString(e));
                    break;
                }
                acc[k] += ms;
            }
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        if (rc) return rc;
        for (int k = 0; k < 5; k++) stage_ms[k] = acc[k] / (double)reps;
        return FCC_OK;
    }

    int run_timed(int64_t n, double *fms, double *ums) override {
        want_timed = true;
        acc_f = acc_u = 0.0;
        acc_n = 0;
        int rc = run(n);
        want_timed = false;
        if (rc) return rc;
        if (fms) *fms = acc_n ? acc_f / acc_n : 0.0;
        if (ums) *ums = acc_n ? acc_u / acc_n : 0.0;
        return FCC_OK;
    }

    int ensure_pipeline() {
        if (snap) return FCC_OK;
        HIPCHK(hipHostMalloc((void **)&snap, sizeof(DevStatus) * (NQ + 1), hipHostMallocDefault));
        HIPCHK(hipHostGetDevicePointer((void **)&snap_d, snap, 0));
        for (auto &e : sev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        for (int q = 0; q < NQ; q++)
            for (auto &e : tev[q]) HIPCHK(hipEventCreate(&e));
        return FCC_OK;
    }

    // k steps, then a snapshot of the status into pinned slot `slot`.
    // Timed runs: one-kernel steps between two events around the chunk's
    // graph launches (inside a graph the step kernels follow each other
    // without a gap, so the interval over k is the step kernel's duration);
    // two-kernel steps launch the chunk's last step eagerly, the same kernels
    // and arguments as the graph, with events around each kernel.
    int issue(int64_t k, int slot) {
        int rc;
        if (!want_timed) {
            if ((rc = replay(k))) return rc;
        } else if (gone) {
            HIPCHK(hipEventRecord(tev[slot][0], stream));
            if ((rc = replay(k))) return rc;
            HIPCHK(hipEventRecord(tev[slot][1], stream));
        } else {
            if ((rc = replay(k - 1))) return rc;
            p.skin_check = (has_contact && skin > 0.0) ? 1 : 0;
            if ((rc = launch_step(tev[slot][0], tev[slot][1], tev[slot][2], tev[slot][3]))) return rc;
        }
        hipLaunchKernelGGL(k_snapshot, 1, 1, 0, stream, d.st, snap_d + slot);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(sev[slot], stream));
        return FCC_OK;
    }

    // Steps per chunk: half the rebuild-interval estimate, 16 to 256. With
    // two chunks queued, a chunk boundary costs the device a status snapshot
    // and a graph launch, while a step frozen by a raised rebuild flag costs
    // a launch whose workgroups return at once; the chunk length does not
    // change results. (The reference, which waits for every chunk, takes a
    // quarter of the estimate, from 4 steps: ext/FourierContourCellsCUDAExt.jl:65-71.)
    int64_t chunk_len() const {
        if (!has_contact) return 256;
        return std::min<int64_t>(256, std::max<int64_t>(16, (int64_t)std::llround(interval / 2)));
    }

    // run! in graph mode, pipelined: the host keeps up to NQ chunks queued
    // and reads each chunk's status snapshot while the next one runs, so the
    // device does not wait for the host between chunks. A raised rebuild flag
    // freezes every later step, those of chunks already queued included,
    // until the host has rebuilt the lists; chunks are queued only for steps
    // still due if every queued step executes, so at most n steps run. The
    // chunk length follows the rebuild-interval estimate of
    // ext/FourierContourCellsCUDAExt.jl:65-71; it does not change results.
    int run(int64_t n) override {
        if (n <= 0) return FCC_OK;
        const bool mirrored = mirror_ok && !member;
        mirror_ok = false;  // set again by a run that ends with the status known
        if (has_contact && skin <= 0.0) return step(n);  // graph_capable (src/simulation.jl:206-207)
        int rc;
        if ((rc = ensure_pipeline())) return rc;
        if (has_contact && !valid && (rc = rebuild())) return rc;
        // The status at the start of the run is one more snapshot (slot NQ),
        // read while the first chunks run (for one-kernel steps written by the
        // slot synchronisation). A rebuild flag raised by the last step before
        // the run freezes them, and the lists are rebuilt before the run goes
        // on (update_neighbors!(sim), src/simulation.jl:132-142). Straight
        // after a run that left no rebuild flag and no pending skin trip (and
        // with nothing in between that changes the status or the state), the
        // status is that run's last snapshot: no start snapshot.
        bool started = mirrored;
        DevStatus st, last = mirror;
        if (!mirrored) {
            if (one_kernel()) {
                if ((rc = sync_slots(snap_d + NQ))) return rc;
            } else {
                hipLaunchKernelGGL(k_snapshot, 1, 1, 0, stream, d.st, snap_d + NQ);
                HIPCHK(hipGetLastError());
            }
            HIPCHK(hipEventRecord(sev[NQ], stream));
        }
        int64_t clock = mirrored ? mirror.clock : 0, done = 0, planned = 0, issued = 0, read = 0;
        int64_t kq[NQ] = {};
        // parity of the buffers holding the state the device has reached
        // (the host's buffers run ahead by the queued one-kernel steps)
        int pact = par;
        bool trip_left = false;  // the last snapshot: a one-kernel step's trip not yet in the flag
        auto account = [&](int64_t executed) {
            for (int64_t q = 0; q < executed; q++) t += cfg.dt;
            nstep += executed;
            done += executed;
            since_rebuild += executed;
        };
        // every queued step after a raised rebuild flag is frozen: let them
        // drain and take the clock they leave (no step of theirs ran)
        auto drain = [&]() -> int {
            HIPCHK(hipMemcpyAsync(&st, d.st, sizeof st, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            last = st;
            if (gone) pact ^= (int)((st.clock - clock) & 1);
            account(st.clock - clock);
            clock = st.clock;
            read = issued;
            planned = done;
            set_parity(pact);
            return FCC_OK;
        };
        while (done < n) {
            while (issued - read < NQ && planned < n) {
                if (!gx[0][0] || gexec_gen != gen) {
                    if (issued > read) break;  // capture with nothing queued
                    if ((rc = capture())) return rc;
                }
                if (builds != gbuilds) {
                    gbuilds = builds;
                    since_rebuild = 0;
                }
                // (a remainder of less than half a chunk joins the chunk)
                const int64_t cl = chunk_len();
                const int64_t k = n - planned < cl + cl / 2 ? n - planned : cl;
                const int slot = (int)(issued % NQ);
                if ((rc = issue(k, slot))) return rc;
                kq[slot] = k;
                planned += k;
                issued++;
            }
            if (!started) {
                HIPCHK(hipEventSynchronize(sev[NQ]));
                clock = snap[NQ].clock;
                last = snap[NQ];
                started = true;
                if (has_contact && snap[NQ].rebuild != 0) {
                    if ((rc = drain()) || (rc = rebuild())) return rc;
                    continue;
                }
            }
            if (read == issued) return fail(FCC_ERR_STATE, "run: no chunk in flight with %lld steps due", (long long)(n - done));
            const int slot = (int)(read % NQ);
            HIPCHK(hipEventSynchronize(sev[slot]));
            const DevStatus s = snap[slot];
            last = s;
            read++;
            const int64_t executed = s.clock - clock;
            clock = s.clock;
            if (gone) pact ^= (int)(executed & 1);
            // a rebuild is due: the flag, or a skin trip of the last one-kernel step
            const bool due = s.rebuild != 0 || (gone && s.tstep[pact] == s.clock - 1);
            trip_left = gone && s.rebuild == 0 && s.tstep[pact] == s.clock - 1;
            if (want_timed && executed == kq[slot]) {  // every step of the chunk ran (none frozen)
                float a = 0.f, b = 0.f;
                HIPCHK(hipEventElapsedTime(&a, tev[slot][0], tev[slot][1]));
                if (!gone) HIPCHK(hipEventElapsedTime(&b, tev[slot][2], tev[slot][3]));
                acc_f += a;
                acc_u += b;
                acc_n += gone ? kq[slot] : 1;  // steps of the chunk, or its eager last step
            }
            account(executed);
            if (due && (rc = drain())) return rc;
            if (due && done < n) {
                if ((rc = rebuild()) || (gone && (rc = sync_slots()))) return rc;
                interval = (interval + (double)since_rebuild) / 2;
                since_rebuild = 0;
                gbuilds = builds;
            } else if ((double)since_rebuild > interval) {
                interval = std::min((double)since_rebuild, 4.0 * 256);
            }
        }
        set_parity(pact);
        if (trip_left) return sync_slots();  // a trip of the last step becomes the rebuild flag
        // (the last snapshot is the status the device is left with)
        mirror = last;
        mirror_ok = last.rebuild == 0 && last.clock == clock && read == issued;
        return FCC_OK;
    }

    int compute_gradients(int frame) override {
        if (has_contact && !valid) return fail(FCC_ERR_STATE, "neighbour list not built (call fcc_update_neighbors)");
        int rc = launch_forces(0);
        if (rc) return rc;
        if (frame) {
            const unsigned g = blocks(N, 256);
            switch (KM) {
            case 5: hipLaunchKernelGGL((k_body_frame<T, 5>), g, 256, 0, stream, p, d); break;
            case 7: hipLaunchKernelGGL((k_body_frame<T, 7>), g, 256, 0, stream, p, d); break;
            case 15: hipLaunchKernelGGL((k_body_frame<T, 15>), g, 256, 0, stream, p, d); break;
            default: hipLaunchKernelGGL((k_body_frame<T, 31>), g, 256, 0, stream, p, d);
            }
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(stream));
        return FCC_OK;
    }

    // ---------------------------------------------------------- transfers
    template <typename U>
    int upload(U *dst, const std::vector<U> &src) {
        HIPCHK(hipMemcpyAsync(dst, src.data(), sizeof(U) * src.size(), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return FCC_OK;
    }
    template <typename U>
    int download(std::vector<U> &dst, const U *src, size_t n) {
        dst.resize(n);
        HIPCHK(hipMemcpyAsync(dst.data(), src, sizeof(U) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return FCC_OK;
    }

    int set_bodies(const fcc_bodies *b) override {
        std::vector<double> x(b->x, b->x + N), y(b->y, b->y + N);
        std::vector<int32_t> ix(N, 0), iy(N, 0), ty(N, 0), pa(N, 0);
        std::vector<T> th(N), r0(N), co((size_t)N * ldc, T(0));
        for (int64_t i = 0; i < N; i++) {
            if (b->imgx) ix[i] = b->imgx[i];
            if (b->imgy) iy[i] = b->imgy[i];
            ty[i] = (b->type ? b->type[i] : 1) - 1;
            pa[i] = (b->parent ? b->parent[i] : (int32_t)(i + 1)) - 1;
            th[i] = (T)b->theta[i];
            r0[i] = (T)b->R0[i];
            for (int c = 0; c < nc; c++) co[(size_t)i * ldc + c] = (T)b->coef[(size_t)i * nc + c];
        }
        int rc;
        if ((rc = upload(d.x, x)) || (rc = upload(d.y, y)) || (rc = upload(d.imgx, ix)) || (rc = upload(d.imgy, iy)) ||
            (rc = upload(d.type, ty)) || (rc = upload(d.parent, pa)) || (rc = upload(d.th, th)) ||
            (rc = upload(d.R0, r0)) || (rc = upload(d.coef, co)))
            return rc;
        if ((rc = launch_init())) return rc;
        HIPCHK(hipStreamSynchronize(stream));
        valid = false;
        return FCC_OK;
    }

    // The time step keeps the lab rows as the shape state; the body-frame
    // coefficients are written to `coef` only for state reads and checkpoints.
    int materialize_coef() {
        const unsigned g = blocks(N, 256);
        switch (KM) {
        case 5: hipLaunchKernelGGL((k_coef_from_lab<T, 5>), g, 256, 0, stream, p, d); break;
        case 7: hipLaunchKernelGGL((k_coef_from_lab<T, 7>), g, 256, 0, stream, p, d); break;
        case 15: hipLaunchKernelGGL((k_coef_from_lab<T, 15>), g, 256, 0, stream, p, d); break;
        default: hipLaunchKernelGGL((k_coef_from_lab<T, 31>), g, 256, 0, stream, p, d);
        }
        HIPCHK(hipGetLastError());
        return FCC_OK;
    }

    int get_bodies(fcc_bodies *b) override {
        std::vector<double> x, y;
        std::vector<int32_t> ix, iy, ty, pa;
        std::vector<T> th, r0, co;
        int rc;
        if ((rc = materialize_coef())) return rc;
        if ((rc = download(x, d.x, N)) || (rc = download(y, d.y, N)) || (rc = download(ix, d.imgx, N)) ||
            (rc = download(iy, d.imgy, N)) || (rc = download(ty, d.type, N)) || (rc = download(pa, d.parent, N)) ||
            (rc = download(th, d.th, N)) || (rc = download(r0, d.R0, N)) || (rc = download(co, d.coef, (size_t)N * ldc)))
            return rc;
        for (int64_t i = 0; i < N; i++) {
            if (b->x) b->x[i] = x[i];
            if (b->y) b->y[i] = y[i];
            if (b->imgx) b->imgx[i] = ix[i];
            if (b->imgy) b->imgy[i] = iy[i];
            if (b->type) b->type[i] = ty[i] + 1;
            if (b->parent) b->parent[i] = pa[i] + 1;
            if (b->theta) b->theta[i] = (double)th[i];
            if (b->R0) b->R0[i] = (double)r0[i];
            if (b->coef)
                for (int c = 0; c < nc; c++) b->coef[(size_t)i * nc + c] = (double)co[(size_t)i * ldc + c];
        }
        return FCC_OK;
    }

    int set_dt(double dt) override {
        cfg.dt = dt;
        p.dt = (T)dt;
        noise_amplitudes();
        gen++;  // dt is baked into the captured graph
        return FCC_OK;
    }

    // √(2·D·dt) of every type in T: (2·D)·dt rounded to T, then the
    // correctly rounded square root, the bits a per-body evaluation on the
    // device gives (IEEE multiply and square root in both places).
    void noise_amplitudes() {
        for (int q = 0; q < FCC_MAX_TYPES; q++) {
            const T a = T(2) * p.Dt[q], b = T(2) * p.Dr[q];
            p.sDt[q] = std::sqrt((T)(a * p.dt));
            p.sDr[q] = std::sqrt((T)(b * p.dt));
        }
    }

    int rows_out(double *out, const T *src) {
        if (!out) return FCC_OK;
        std::vector<T> h;
        int rc = download(h, src, (size_t)N * ldc);
        if (rc) return rc;
        for (int64_t i = 0; i < N; i++)
            for (int c = 0; c < nc; c++) out[(size_t)i * nc + c] = (double)h[(size_t)i * ldc + c];
        return FCC_OK;
    }
    int vec_out(double *out, const T *src, size_t n) {
        if (!out) return FCC_OK;
        std::vector<T> h;
        int rc = download(h, src, n);
        if (rc) return rc;
        for (size_t i = 0; i < n; i++) out[i] = (double)h[i];
        return FCC_OK;
    }

    int get_gradients(double *gx, double *gy, double *gth, double *glab, double *gcoef, double *E) override {
        int rc;
        if ((rc = vec_out(gx, d.gx, N)) || (rc = vec_out(gy, d.gy, N)) || (rc = vec_out(gth, d.gth, N)) ||
            (rc = rows_out(glab, d.glab)) || (rc = rows_out(gcoef, d.gcoef)) || (rc = vec_out(E, d.Ebody, N)))
            return rc;
        return FCC_OK;
    }

    int get_contour_gradients(double *P, double *GR, double *GD, double *GDD) override {
        if (!debug) return fail(FCC_ERR_STATE, "contour gradients need FCC_FLAG_DEBUG_BUFFERS");
        const size_t m = (size_t)N * nth;
        int rc;
        if ((rc = vec_out(P, d.P, m)) || (rc = vec_out(GR, d.GR, m)) || (rc = vec_out(GD, d.GD, m)) ||
            (rc = vec_out(GDD, d.GDD, m)))
            return rc;
        return FCC_OK;
    }

    int launch_geometry(T *R, T *dR, T *ddR, T *per, T *area) {
        const unsigned g = blocks(N, 256);
        switch (KM) {
        case 5: hipLaunchKernelGGL((k_geometry<T, 5>), g, 256, 0, stream, p, d, R, dR, ddR, per, area); break;
        case 7: hipLaunchKernelGGL((k_geometry<T, 7>), g, 256, 0, stream, p, d, R, dR, ddR, per, area); break;
        case 15: hipLaunchKernelGGL((k_geometry<T, 15>), g, 256, 0, stream, p, d, R, dR, ddR, per, area); break;
        default: hipLaunchKernelGGL((k_geometry<T, 31>), g, 256, 0, stream, p, d, R, dR, ddR, per, area);
        }
        HIPCHK(hipGetLastError());
        return FCC_OK;
    }

    int get_geometry(double *lab, double *rmax, double *R, double *dR, double *ddR) override {
        int rc;
        if ((rc = rows_out(lab, d.lab)) || (rc = vec_out(rmax, d.rmax, N))) return rc;
        if (R || dR || ddR) {
            const size_t m = (size_t)N * nth;
            T *tR, *tD, *tDD;
            if ((rc = alloc(&tR, m)) || (rc = alloc(&tD, m)) || (rc = alloc(&tDD, m))) return rc;
            if (!(rc = launch_geometry(tR, tD, tDD, nullptr, nullptr))) {
                if (!(rc = vec_out(R, tR, m)) && !(rc = vec_out(dR, tD, m))) rc = vec_out(ddR, tDD, m);
            }
            release(tR);
            release(tD);
            release(tDD);
            return rc;
        }
        return FCC_OK;
    }

    int shape_index(double *per, double *area, double *pidx) override {
        T *tp, *ta;
        int rc;
        if ((rc = alloc(&tp, N)) || (rc = alloc(&ta, N))) return rc;
        std::vector<T> hp, ha;
        if (!(rc = launch_geometry(nullptr, nullptr, nullptr, tp, ta)) && !(rc = download(hp, tp, N)))
            rc = download(ha, ta, N);
        release(tp);
        release(ta);
        if (rc) return rc;
        for (int64_t i = 0; i < N; i++) {
            if (per) per[i] = (double)hp[i];
            if (area) area[i] = (double)ha[i];
            if (pidx) pidx[i] = (double)(hp[i] / std::sqrt(ha[i]));
        }
        return FCC_OK;
    }

    int get_celllist(int32_t *ox, int32_t *oy, int32_t *cid, int32_t *pm, int32_t *st, int32_t *sp) override {
        if (ox) *ox = ncx;
        if (oy) *oy = ncy;
        std::vector<int32_t> a, b, s, c;
        int rc;
        if ((rc = download(a, cellid, N)) || (rc = download(b, perm, N)) || (rc = download(s, cstart, nce)) ||
            (rc = download(c, ccnt, nce)))
            return rc;
        for (int64_t i = 0; i < N; i++) {
            if (cid) cid[i] = a[i] + 1;
            if (pm) pm[i] = b[i] + 1;
        }
        for (int64_t q = 0; q < nce; q++) {
            const bool empty = c[q] == 0;
            if (st) st[q] = empty ? 1 : s[q] + 1;
            if (sp) sp[q] = empty ? 0 : s[q] + c[q];
        }
        return FCC_OK;
    }

    int get_neighbors(int32_t *cap, int32_t *count, int32_t *idx) override {
        if (cap) *cap = d.cap;
        std::vector<int32_t> cn, ni;
        int rc;
        if ((rc = download(cn, d.ncount, N))) return rc;
        if (count) std::copy(cn.begin(), cn.end(), count);
        if (idx) {
            if ((rc = download(ni, d.nidx, (size_t)d.cap * N))) return rc;
            for (int64_t i = 0; i < N; i++)
                for (int32_t q = 0; q < d.cap; q++)
                    idx[(size_t)i * d.cap + q] = q < cn[i] ? ni[(size_t)q * N + i] + 1 : 0;
        }
        return FCC_OK;
    }

    int set_neighbor_capacity(int32_t cap) override {
        if (cap < 1) return fail(FCC_ERR_ARGUMENT, "capacity must be positive");
        int32_t *ni = nullptr;
        int rc = alloc(&ni, (size_t)cap * N);
        if (rc) return rc;
        HIPCHK(hipMemsetAsync(ni, 0, sizeof(int32_t) * cap * N, stream));
        HIPCHK(hipStreamSynchronize(stream));
        release(d.nidx);
        d.nidx = ni;
        d.cap = cap;
        valid = false;
        gen++;
        return FCC_OK;
    }

    // Device regions a step writes (state, geometry, gradients, status).
    std::vector<std::pair<void *, size_t>> step_regions() {
        const size_t n = (size_t)N, rows = n * (size_t)ldc;
        return {{d.x, 8 * n},       {d.y, 8 * n},           {d.imgx, 4 * n},
                {d.imgy, 4 * n},    {d.th, sizeof(T) * n},  {d.coef, sizeof(T) * rows},
                {d.lab, sizeof(T) * rows}, {d.rmax, sizeof(T) * n}, {d.rmin, sizeof(T) * n},
                {d.glab, sizeof(T) * rows}, {d.gx, sizeof(T) * n}, {d.gy, sizeof(T) * n},
                {d.gth, sizeof(T) * n},     {d.Ebody, sizeof(T) * n}, {d.st, sizeof(DevStatus)}};
    }

    // Two series between HIP events: `reps` back-to-back forces launches,
    // then `reps` steps (forces + update kernels, which advance the state:
    // an update launched on its own again and again would integrate stale
    // forces and time a state the model never reaches). Forces = series 1 /
    // reps, update = (series 2 - series 1) / reps. Each series is a captured
    // graph, as in fcc_run; the forces series is first replayed untimed
    // (about 400 launches: an idle GPU runs at lower clocks for the first
    // milliseconds of work).
    // The state is saved before and restored after, so the simulation is
    // unchanged.
    int time_kernels(int64_t reps, double *fms, double *ums) override {
        if (has_contact && !valid) return fail(FCC_ERR_STATE, "neighbour list not built (call fcc_update_neighbors)");
        const bool one = one_kernel();
        if (one) {
            int rc = ensure_next();
            if (rc) return rc;
        }
        if (reps < 1) return fail(FCC_ERR_ARGUMENT, "reps must be positive");
        const auto regions = step_regions();
        size_t bytes = 0;
        for (const auto &q : regions) bytes += (q.second + 255) / 256 * 256;
        char *snap = nullptr;
        HIPCHK(hipMalloc(&snap, bytes));
        hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
        hipGraphExec_t gx[2] = {nullptr, nullptr};
        int rc = FCC_OK;
        auto copy = [&](bool save) -> int {
            size_t off = 0;
            for (const auto &q : regions) {
                if (save)
                    HIPCHK(hipMemcpyAsync(snap + off, q.first, q.second, hipMemcpyDeviceToDevice, stream));
                else
                    HIPCHK(hipMemcpyAsync(q.first, snap + off, q.second, hipMemcpyDeviceToDevice, stream));
                off += (q.second + 255) / 256 * 256;
            }
            return FCC_OK;
        };
        // series s: reps forces launches (s = 0) or reps steps (s = 1)
        // (one-kernel steps: without the freeze check, the host's buffers
        // restored after the capture)
        auto capture_series = [&](int s, hipGraphExec_t *out) -> int {
            std::lock_guard<std::mutex> lk(g_capture_lock);
            hipGraph_t graph;
            const int par0 = par;
            HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            int r = FCC_OK;
            for (int64_t q = 0; q < reps && !r; q++) {
                if (s == 1 && one) {
                    if (!(r = launch_one_kernel(4, nullptr, nullptr))) flip();
                } else if (!(r = launch_forces(0)) && s == 1) {
                    r = launch_update();
                }
            }
            hipError_t e = hipStreamEndCapture(stream, &graph);
            set_parity(par0);
            if (r) return r;
            HIPCHK(e);
            e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIPCHK(e);
            return FCC_OK;
        };
        for (auto &e : ev)
            if (!rc && hipEventCreate(&e) != hipSuccess) rc = fail(FCC_ERR_DEVICE, "hipEventCreate failed");
        p.skin_check = (has_contact && skin > 0.0) ? 1 : 0;
        if (!rc) rc = capture_series(0, &gx[0]);
        if (!rc) rc = capture_series(1, &gx[1]);
        if (!rc) rc = copy(true);
        if (!rc) {
            hipLaunchKernelGGL(k_clear_status, 1, 1, 0, stream, d.st);  // update = 0: the update kernel runs
            // untimed replays, about 20 ms of work at N = 262144 (at least one)
            hipError_t e = hipSuccess;
            for (int64_t w = 0; w < std::max<int64_t>(1, 400 / reps) && e == hipSuccess; w++)
                e = hipGraphLaunch(gx[0], stream);
            if (e == hipSuccess) e = hipEventRecord(ev[0], stream);
            if (e == hipSuccess) e = hipGraphLaunch(gx[0], stream);
            if (e == hipSuccess) e = hipEventRecord(ev[1], stream);
            if (e == hipSuccess) e = hipGraphLaunch(gx[1], stream);
            if (e == hipSuccess) e = hipEventRecord(ev[2], stream);
            if (e == hipSuccess) e = hipEventSynchronize(ev[2]);
            float a = 0.f, b = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&a, ev[0], ev[1]);
            if (e == hipSuccess) e = hipEventElapsedTime(&b, ev[1], ev[2]);
            if (e != hipSuccess) rc = fail(FCC_ERR_DEVICE, "kernel timing: %s", hipGetErrorString(e));
            if (!rc) {
                if (fms) *fms = a / (double)reps;
                if (ums) *ums = (b - a) / (double)reps;
            }
            const int rc2 = copy(false);
            if (!rc) rc = rc2;
        }
        (void)hipStreamSynchronize(stream);
        for (hipGraphExec_t g : gx)
            if (g) (void)hipGraphExecDestroy(g);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        (void)hipFree(snap);
        return rc;
    }

    int time_step_kernels(int64_t steps, double *fms, double *ums) override {
        hipEvent_t e0, e1, e2, e3;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventCreate(&e2));
        HIPCHK(hipEventCreate(&e3));
        double sf = 0.0, su = 0.0;
        int rc = FCC_OK;
        for (int64_t q = 0; q < steps && !rc; q++) {
            if (has_contact && (rc = update_neighbors(0))) break;
            p.skin_check = (has_contact && valid && skin > 0.0) ? 1 : 0;
            hipError_t e = hipSuccess;
            const bool one = one_kernel();
            if (one && ((rc = ensure_next()) || (rc = sync_slots()))) break;
            rc = launch_step(e0, e1, one ? nullptr : e2, one ? nullptr : e3);
            if (!rc && one) rc = sync_slots();
            if (!rc) e = hipEventSynchronize(one ? e1 : e3);
            float a = 0.f, b = 0.f;
            if (!rc && e == hipSuccess) e = hipEventElapsedTime(&a, e0, e1);
            if (!rc && !one && e == hipSuccess) e = hipEventElapsedTime(&b, e2, e3);
            if (!rc && e != hipSuccess) rc = fail(FCC_ERR_DEVICE, "event timing: %s", hipGetErrorString(e));
            sf += a;
            su += b;
            t += cfg.dt;
            nstep++;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipEventDestroy(e2);
        (void)hipEventDestroy(e3);
        if (rc) return rc;
        if (fms) *fms = steps > 0 ? sf / (double)steps : 0.0;
        if (ums) *ums = steps > 0 ? su / (double)steps : 0.0;
        return FCC_OK;
    }

    // BondList(bonds, N) (src/energies/bonds.jl:54-72): checks, then each
    // bond (i, j) appended to the partner lists of i and of j in the order
    // given; stored slot-major for coalesced reads.
    int set_bonds(int64_t n, const int32_t *pairs) override {
        if (n < 0 || (n > 0 && !pairs)) return fail(FCC_ERR_ARGUMENT, "bad bond arguments");
        std::vector<int32_t> count(N, 0);
        std::vector<int64_t> key(n);
        for (int64_t q = 0; q < n; q++) {
            const int32_t i = pairs[2 * q], j = pairs[2 * q + 1];
            if (i < 1 || i > N || j < 1 || j > N)
                return fail(FCC_ERR_ARGUMENT, "bond (%d, %d) refers to a body outside 1:%lld", i, j, (long long)N);
            if (i == j) return fail(FCC_ERR_ARGUMENT, "bond (%d, %d) connects a body to itself", i, j);
            key[q] = (int64_t)std::min(i, j) * (N + 1) + std::max(i, j);
            count[i - 1]++;
            count[j - 1]++;
        }
        std::sort(key.begin(), key.end());
        for (int64_t q = 1; q < n; q++)
            if (key[q] == key[q - 1])
                return fail(FCC_ERR_ARGUMENT, "duplicate bond (%lld, %lld)", (long long)(key[q] / (N + 1)),
                            (long long)(key[q] % (N + 1)));
        int32_t cap = 1;
        for (int32_t c : count) cap = std::max(cap, c);
        std::vector<int32_t> idx((size_t)cap * N, 0);
        std::fill(count.begin(), count.end(), 0);
        for (int64_t q = 0; q < n; q++) {
            const int32_t a = pairs[2 * q] - 1, b = pairs[2 * q + 1] - 1;
            idx[(size_t)count[a]++ * N + a] = b;
            idx[(size_t)count[b]++ * N + b] = a;
        }
        HIPCHK(hipStreamSynchronize(stream));  // no step in flight reads the old list
        release(d.bidx);
        release(d.bcount);
        d.bidx = d.bcount = nullptr;
        nbonds = n;
        if (n > 0) {
            int rc;
            if ((rc = alloc(&d.bidx, idx.size())) || (rc = alloc(&d.bcount, (size_t)N)) || (rc = upload(d.bidx, idx)) ||
                (rc = upload(d.bcount, count)))
                return rc;
        }
        has_bonds = p.nbondt > 0 && n > 0;
        gen++;  // the captured step gains or loses the bond kernel
        return FCC_OK;
    }

    // ---------------------------------------------------- contour frames
    int launch_contours(double *X, double *Y) {
        const unsigned g = blocks(N, 256);
        switch (KM) {
        case 5: hipLaunchKernelGGL((k_contours<T, 5>), g, 256, 0, stream, p, d, X, Y); break;
        case 7: hipLaunchKernelGGL((k_contours<T, 7>), g, 256, 0, stream, p, d, X, Y); break;
        case 15: hipLaunchKernelGGL((k_contours<T, 15>), g, 256, 0, stream, p, d, X, Y); break;
        default: hipLaunchKernelGGL((k_contours<T, 31>), g, 256, 0, stream, p, d, X, Y);
        }
        HIPCHK(hipGetLastError());
        return FCC_OK;
    }

    int contours(double *X, double *Y) override {
        if (!X || !Y) return fail(FCC_ERR_ARGUMENT, "null output arrays");
        const size_t m = (size_t)N * nth;
        double *dX, *dY;
        int rc;
        if ((rc = alloc(&dX, m)) || (rc = alloc(&dY, m))) return rc;
        if (!(rc = launch_contours(dX, dY))) {
            hipError_t e = hipMemcpyAsync(X, dX, sizeof(double) * m, hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipMemcpyAsync(Y, dY, sizeof(double) * m, hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e != hipSuccess) rc = fail(FCC_ERR_DEVICE, "contour copy: %s", hipGetErrorString(e));
        }
        release(dX);
        release(dY);
        return rc;
    }

    // Frames for live views: the contour kernel runs in stream order into a
    // device staging buffer; a copy stream moves it to pinned host memory
    // while the simulation stream goes on with the next steps.
    struct Frame {
        double *dX = nullptr, *dY = nullptr;  // device staging (dX holds X, Y and the clock)
        double *hX = nullptr;                 // pinned host copy, same layout
        hipStream_t copy = nullptr;
        hipEvent_t ready = nullptr, done = nullptr;
        bool pending = false, used = false;
    } fr;

    int frame_request() override {
        const size_t m = (size_t)N * nth;
        if (!fr.copy) {
            HIPCHK(hipStreamCreateWithFlags(&fr.copy, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&fr.ready, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&fr.done, hipEventDisableTiming));
            int rc;
            if ((rc = alloc(&fr.dX, 2 * m + 1))) return rc;
            fr.dY = fr.dX + m;
            HIPCHK(hipHostMalloc((void **)&fr.hX, sizeof(double) * (2 * m + 1), hipHostMallocDefault));
        }
        // the staging buffer is rewritten only after the last frame's copy
        if (fr.used) HIPCHK(hipStreamWaitEvent(stream, fr.done, 0));
        int rc = launch_contours(fr.dX, fr.dY);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(fr.dX + 2 * m, &d.st->clock, sizeof(int64_t), hipMemcpyDeviceToDevice, stream));
        HIPCHK(hipEventRecord(fr.ready, stream));
        HIPCHK(hipStreamWaitEvent(fr.copy, fr.ready, 0));
        HIPCHK(hipMemcpyAsync(fr.hX, fr.dX, sizeof(double) * (2 * m + 1), hipMemcpyDeviceToHost, fr.copy));
        HIPCHK(hipEventRecord(fr.done, fr.copy));
        fr.pending = fr.used = true;
        return FCC_OK;
    }

    int frame_read(double *X, double *Y, int64_t *clock) override {
        if (!fr.pending) return fail(FCC_ERR_STATE, "no frame requested (fcc_frame_request)");
        HIPCHK(hipEventSynchronize(fr.done));
        const size_t m = (size_t)N * nth;
        if (X) std::memcpy(X, fr.hX, sizeof(double) * m);
        if (Y) std::memcpy(Y, fr.hX + m, sizeof(double) * m);
        if (clock) std::memcpy(clock, fr.hX + 2 * m, sizeof(int64_t));
        fr.pending = false;
        return FCC_OK;
    }

    // ---------------------------------------------------- checkpoints
    // Everything a continued run reads: the body state, the lab frame and
    // bounds of the last update, the Verlet reference and the neighbour list
    // (its order fixes the summation order of the contact sums), the device
    // clock (Philox draw index) and rebuild flag, and the host counters.
    struct CkptHeader {
        char magic[8];
        int32_t version, precision;
        int64_t N;
        int32_t ncoef, ldc, nth, cap;
        int64_t clock, step, builds;
        double t, dt, Lx, Ly, skin;
        uint64_t seed;
        int32_t rebuild, valid, maxcount, ntypes;
    };
    static constexpr char CKPT_MAGIC[8] = {'F', 'C', 'C', 'C', 'K', 'P', 'T', '1'};

    size_t ckpt_bytes(int32_t cap) const {
        const size_t n = (size_t)N;
        return sizeof(CkptHeader) + n * (2 * sizeof(double) + 4 * sizeof(int32_t) + 4 * sizeof(T)) +
               2 * n * ldc * sizeof(T) + n * (2 * sizeof(double) + sizeof(T)) + (size_t)cap * n * sizeof(int32_t) +
               n * sizeof(int32_t);
    }

    // (device pointer, bytes) of every array in the checkpoint, in order
    std::vector<std::pair<void *, size_t>> ckpt_arrays() {
        const size_t n = (size_t)N;
        return {{d.x, n * 8},           {d.y, n * 8},          {d.imgx, n * 4},           {d.imgy, n * 4},
                {d.type, n * 4},        {d.parent, n * 4},     {d.th, n * sizeof(T)},     {d.R0, n * sizeof(T)},
                {d.rmax, n * sizeof(T)}, {d.rmin, n * sizeof(T)}, {d.coef, n * ldc * sizeof(T)},
                {d.lab, n * ldc * sizeof(T)}, {d.x0, n * 8},    {d.y0, n * 8},             {d.rmax0, n * sizeof(T)},
                {d.nidx, (size_t)d.cap * n * 4}, {d.ncount, n * 4}};
    }

    int checkpoint(void *buf, int64_t bytes, int64_t *need) override {
        const size_t total = ckpt_bytes(d.cap);
        if (need) *need = (int64_t)total;
        if (!buf) return FCC_OK;
        if (bytes < (int64_t)total) return fail(FCC_ERR_DIMENSION, "checkpoint needs %lld bytes", (long long)total);
        DevStatus st;
        HIPCHK(hipMemcpyAsync(&st, d.st, sizeof st, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        CkptHeader h{};
        std::memcpy(h.magic, CKPT_MAGIC, 8);
        h.version = 1;
        h.precision = cfg.precision;
        h.N = N;
        h.ncoef = nc;
        h.ldc = ldc;
        h.nth = nth;
        h.cap = d.cap;
        h.clock = st.clock;
        h.step = nstep;
        h.builds = builds;
        h.t = t;
        h.dt = cfg.dt;
        h.Lx = cfg.Lx;
        h.Ly = cfg.Ly;
        h.skin = cfg.skin;
        h.seed = cfg.seed;
        h.rebuild = st.rebuild;
        h.valid = valid ? 1 : 0;
        h.maxcount = maxcount;
        h.ntypes = cfg.n_types;
        char *out = (char *)buf;
        std::memcpy(out, &h, sizeof h);
        size_t off = sizeof h;
        if (int rc = materialize_coef()) return rc;
        for (auto &a : ckpt_arrays()) {
            HIPCHK(hipMemcpyAsync(out + off, a.first, a.second, hipMemcpyDeviceToHost, stream));
            off += a.second;
        }
        HIPCHK(hipStreamSynchronize(stream));
        return FCC_OK;
    }

    int restore(const void *buf, int64_t bytes) override {
        if (!buf || bytes < (int64_t)sizeof(CkptHeader)) return fail(FCC_ERR_DIMENSION, "checkpoint too short");
        CkptHeader h;
        std::memcpy(&h, buf, sizeof h);
        if (std::memcmp(h.magic, CKPT_MAGIC, 8) != 0 || h.version != 1)
            return fail(FCC_ERR_ARGUMENT, "not a checkpoint of this library");
        if (h.precision != cfg.precision || h.N != N || h.ncoef != nc || h.ldc != ldc || h.nth != nth)
            return fail(FCC_ERR_DIMENSION, "checkpoint of another precision or size");
        if (h.dt != cfg.dt || h.Lx != cfg.Lx || h.Ly != cfg.Ly || h.skin != cfg.skin || h.seed != cfg.seed ||
            h.ntypes != cfg.n_types)
            return fail(FCC_ERR_ARGUMENT, "checkpoint of another configuration (dt, box, skin, seed or types)");
        if (h.cap < 1 || bytes < (int64_t)ckpt_bytes(h.cap)) return fail(FCC_ERR_DIMENSION, "checkpoint truncated");
        const char *in = (const char *)buf;
        // body types and parents fix the children lists built at creation
        {
            const size_t n = (size_t)N, off = sizeof h + n * (8 + 8 + 4 + 4);
            std::vector<int32_t> ty, pa;
            int rc;
            if ((rc = download(ty, d.type, n)) || (rc = download(pa, d.parent, n))) return rc;
            if (std::memcmp(ty.data(), in + off, n * 4) != 0 || std::memcmp(pa.data(), in + off + n * 4, n * 4) != 0)
                return fail(FCC_ERR_ARGUMENT, "checkpoint with other body types or parents");
        }
        HIPCHK(hipStreamSynchronize(stream));
        if (h.cap != d.cap) {
            int32_t *ni = nullptr;
            int rc = alloc(&ni, (size_t)h.cap * N);
            if (rc) return rc;
            release(d.nidx);
            d.nidx = ni;
            d.cap = h.cap;
        }
        size_t off = sizeof h;
        for (auto &a : ckpt_arrays()) {
            HIPCHK(hipMemcpyAsync(a.first, in + off, a.second, hipMemcpyHostToDevice, stream));
            off += a.second;
        }
        DevStatus st;
        HIPCHK(hipMemcpyAsync(&st, d.st, sizeof st, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        st.rebuild = h.rebuild;
        st.update = 0;
        st.overflow = 0;
        st.stop = 0;
        st.clock = h.clock;
        st.target = INT64_MAX;
        st.tstep[0] = st.tstep[1] = INT64_MIN;  // a pending skin trip is in the saved rebuild flag
        HIPCHK(hipMemcpyAsync(d.st, &st, sizeof st, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        nstep = h.step;
        builds = h.builds;
        t = h.t;
        valid = h.valid != 0;
        maxcount = h.maxcount;
        gen++;  // kernel choice and list buffers may differ from the captured step
        return FCC_OK;
    }

    // ---------------------------------------------------- observables
    // Device buffers of the current observation (fcc_observer_*).
    struct Observer {
        bool on = false;
        fcc_observer_config cfg{};
        int64_t nblk = 0, n = 0;  // workgroups of k_observe; samples taken
        double *x0 = nullptr, *y0 = nullptr, *part = nullptr, *msd = nullptr, *pm = nullptr, *p2 = nullptr;
        unsigned long long *hist = nullptr;
        int64_t *clk = nullptr;
    } obs;

    void observer_free() {
        release(obs.x0);
        release(obs.y0);
        release(obs.part);
        release(obs.msd);
        release(obs.pm);
        release(obs.p2);
        release(obs.hist);
        release(obs.clk);
        obs = Observer{};
    }

    int observer_start(const fcc_observer_config *c) override {
        HIPCHK(hipStreamSynchronize(stream));  // no sample of the old observation is in flight
        observer_free();
        const int nt = cfg.n_types;
        const size_t ms = (size_t)c->max_samples, nh = (size_t)nt * (c->nbins + 2);
        obs.cfg = *c;
        obs.nblk = blocks(N, OBS_BLOCK);
        int rc;
        if ((rc = alloc(&obs.x0, N)) || (rc = alloc(&obs.y0, N)) || (rc = alloc(&obs.part, obs.nblk * nt * OBS_NV)) ||
            (rc = alloc(&obs.msd, ms * nt)) || (rc = alloc(&obs.pm, ms * nt)) || (rc = alloc(&obs.p2, ms * nt)) ||
            (rc = alloc(&obs.hist, nh)) || (rc = alloc(&obs.clk, ms))) {
            observer_free();
            return rc;
        }
        HIPCHK(hipMemsetAsync(obs.hist, 0, sizeof(unsigned long long) * nh, stream));
        hipLaunchKernelGGL(k_unwrapped, blocks(N, 256), 256, 0, stream, N, d.x, d.y, d.imgx, d.imgy, p.Lx, p.Ly,
                           obs.x0, obs.y0);
        HIPCHK(hipGetLastError());
        obs.on = true;
        return FCC_OK;
    }

    int observe() override {
        if (!obs.on) return fail(FCC_ERR_STATE, "no observation started (fcc_observer_start)");
        if (obs.n >= obs.cfg.max_samples)
            return fail(FCC_ERR_STATE, "the observation holds max_samples = %d samples", obs.cfg.max_samples);
        const int nt = cfg.n_types;
        const ObsArgs o{obs.x0, obs.y0, obs.cfg.pmin, (double)obs.cfg.nbins / (obs.cfg.pmax - obs.cfg.pmin),
                        obs.cfg.nbins, nt, obs.part, obs.hist};
        const size_t sh = sizeof(uint32_t) * nt * (obs.cfg.nbins + 2);
        const unsigned g = (unsigned)obs.nblk;
        switch (KM) {
        case 5: hipLaunchKernelGGL((k_observe<T, 5>), g, OBS_BLOCK, sh, stream, p, d, o); break;
        case 7: hipLaunchKernelGGL((k_observe<T, 7>), g, OBS_BLOCK, sh, stream, p, d, o); break;
        case 15: hipLaunchKernelGGL((k_observe<T, 15>), g, OBS_BLOCK, sh, stream, p, d, o); break;
        default: hipLaunchKernelGGL((k_observe<T, 31>), g, OBS_BLOCK, sh, stream, p, d, o);
        }
        hipLaunchKernelGGL(k_observe_finish, 1, OBS_BLOCK, 0, stream, obs.part, obs.nblk, nt, obs.n, d.st, obs.msd,
                           obs.pm, obs.p2, obs.clk);
        HIPCHK(hipGetLastError());
        obs.n++;
        return FCC_OK;
    }

    int observer_read(int32_t *ns, int64_t *clock, double *msd, double *pm, double *p2, uint64_t *hist) override {
        if (!obs.on) return fail(FCC_ERR_STATE, "no observation started (fcc_observer_start)");
        const size_t nt = (size_t)cfg.n_types, n = (size_t)obs.n;
        if (ns) *ns = (int32_t)obs.n;
        if (clock && n) HIPCHK(hipMemcpyAsync(clock, obs.clk, sizeof(int64_t) * n, hipMemcpyDeviceToHost, stream));
        if (msd && n) HIPCHK(hipMemcpyAsync(msd, obs.msd, sizeof(double) * n * nt, hipMemcpyDeviceToHost, stream));
        if (pm && n) HIPCHK(hipMemcpyAsync(pm, obs.pm, sizeof(double) * n * nt, hipMemcpyDeviceToHost, stream));
        if (p2 && n) HIPCHK(hipMemcpyAsync(p2, obs.p2, sizeof(double) * n * nt, hipMemcpyDeviceToHost, stream));
        if (hist)
            HIPCHK(hipMemcpyAsync(hist, obs.hist, sizeof(uint64_t) * nt * (obs.cfg.nbins + 2), hipMemcpyDeviceToHost,
                                  stream));
        HIPCHK(hipStreamSynchronize(stream));
        return FCC_OK;
    }

    int get_counters(int64_t *st, double *tt, int64_t *clock, int64_t *b) override {
        DevStatus s;
        HIPCHK(hipMemcpyAsync(&s, d.st, sizeof s, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        if (st) *st = nstep;
        if (tt) *tt = t;
        if (clock) *clock = s.clock;
        if (b) *b = builds;
        return FCC_OK;
    }
};

#include "fcc_ensemble.hpp"

// ------------------------------------------------------------------ C ABI

extern "C" {

int fcc_abi_version(void) { return FCC_ABI_VERSION; }
const char *fcc_last_error(void) { return g_err; }

int fcc_create(const fcc_config *cfg, const fcc_bodies *bodies, fcc_sim **out) {
    if (!out) return fail(FCC_ERR_ARGUMENT, "null output handle");
    *out = nullptr;
    int rc = validate(cfg, bodies);
    if (rc) return rc;
    fcc_sim *s = nullptr;
    if (cfg->precision == FCC_FP64) {
        auto *q = new Sim<double>();
        rc = q->init(cfg, bodies);
        s = q;
    } else {
        auto *q = new Sim<float>();
        rc = q->init(cfg, bodies);
        s = q;
    }
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return FCC_OK;
}

int fcc_destroy(fcc_sim *s) {
    if (s && s->borrowed) return fail(FCC_ERR_STATE, "replica handle belongs to an ensemble (fcc_ensemble_destroy)");
    delete s;
    return FCC_OK;
}

#define FCC_HANDLE(s) \
    if (!(s)) return fail(FCC_ERR_ARGUMENT, "null simulation handle")
// entry points that may change the device status or the state
#define FCC_HANDLE_MUT(s) \
    FCC_HANDLE(s);        \
    (s)->mirror_ok = false

int fcc_set_bodies(fcc_sim *s, const fcc_bodies *b) { FCC_HANDLE_MUT(s); return s->set_bodies(b); }
int fcc_get_bodies(fcc_sim *s, fcc_bodies *b) { FCC_HANDLE(s); return s->get_bodies(b); }
int fcc_set_dt(fcc_sim *s, double dt) { FCC_HANDLE_MUT(s); return s->set_dt(dt); }
int fcc_set_bonds(fcc_sim *s, int64_t n, const int32_t *pairs) { FCC_HANDLE_MUT(s); return s->set_bonds(n, pairs); }
int fcc_step(fcc_sim *s, int64_t n) { FCC_HANDLE_MUT(s); return s->step(n); }
int fcc_run(fcc_sim *s, int64_t n) { FCC_HANDLE(s); return s->run(n); }
int fcc_update_neighbors(fcc_sim *s, int32_t force) { FCC_HANDLE_MUT(s); return s->update_neighbors(force); }
int fcc_compute_gradients(fcc_sim *s, int32_t frame) { FCC_HANDLE_MUT(s); return s->compute_gradients(frame); }
int fcc_get_gradients(fcc_sim *s, double *gx, double *gy, double *gth, double *glab, double *gcoef, double *E) {
    FCC_HANDLE(s);
    return s->get_gradients(gx, gy, gth, glab, gcoef, E);
}
int fcc_get_contour_gradients(fcc_sim *s, double *P, double *GR, double *GD, double *GDD) {
    FCC_HANDLE(s);
    return s->get_contour_gradients(P, GR, GD, GDD);
}
int fcc_get_geometry(fcc_sim *s, double *lab, double *rmax, double *R, double *dR, double *ddR) {
    FCC_HANDLE(s);
    return s->get_geometry(lab, rmax, R, dR, ddR);
}
int fcc_shape_index(fcc_sim *s, double *per, double *area, double *p) { FCC_HANDLE(s); return s->shape_index(per, area, p); }
int fcc_get_celllist(fcc_sim *s, int32_t *ncx, int32_t *ncy, int32_t *cellid, int32_t *perm, int32_t *start,
                     int32_t *stop) {
    FCC_HANDLE(s);
    return s->get_celllist(ncx, ncy, cellid, perm, start, stop);
}
int fcc_get_neighbors(fcc_sim *s, int32_t *cap, int32_t *count, int32_t *idx) {
    FCC_HANDLE(s);
    return s->get_neighbors(cap, count, idx);
}
int fcc_set_neighbor_capacity(fcc_sim *s, int32_t cap) { FCC_HANDLE_MUT(s); return s->set_neighbor_capacity(cap); }
int fcc_get_counters(fcc_sim *s, int64_t *step, double *t, int64_t *clock, int64_t *builds) {
    FCC_HANDLE(s);
    return s->get_counters(step, t, clock, builds);
}

int fcc_time_step_kernels(fcc_sim *s, int64_t steps, double *fms, double *ums) {
    FCC_HANDLE_MUT(s);
    if (steps < 0) return fail(FCC_ERR_ARGUMENT, "steps must be non-negative");
    return s->time_step_kernels(steps, fms, ums);
}

int fcc_time_kernels(fcc_sim *s, int64_t reps, double *fms, double *ums) {
    FCC_HANDLE(s);
    if (reps < 1) return fail(FCC_ERR_ARGUMENT, "reps must be positive");
    return s->time_kernels(reps, fms, ums);
}

int fcc_time_rebuild(fcc_sim *s, int64_t reps, double *stage_ms) {
    FCC_HANDLE_MUT(s);
    if (reps < 1 || !stage_ms) return fail(FCC_ERR_ARGUMENT, "need reps >= 1 and an output array");
    return s->time_rebuild(reps, stage_ms);
}

int fcc_run_timed(fcc_sim *s, int64_t n, double *fms, double *ums) {
    FCC_HANDLE(s);
    if (n < 0) return fail(FCC_ERR_ARGUMENT, "steps must be non-negative");
    return s->run_timed(n, fms, ums);
}

static int check_observer(const fcc_observer_config *c) {
    if (!c) return fail(FCC_ERR_ARGUMENT, "null observer configuration");
    if (c->nbins < 1 || c->nbins > FCC_OBS_MAX_BINS) return fail(FCC_ERR_ARGUMENT, "nbins must be in 1:%d", FCC_OBS_MAX_BINS);
    if (c->max_samples < 1) return fail(FCC_ERR_ARGUMENT, "max_samples must be positive");
    if (!(c->pmin < c->pmax) || !std::isfinite(c->pmin) || !std::isfinite(c->pmax))
        return fail(FCC_ERR_ARGUMENT, "the histogram range needs finite pmin < pmax");
    return FCC_OK;
}

int fcc_observer_start(fcc_sim *s, const fcc_observer_config *c) {
    FCC_HANDLE_MUT(s);
    int rc = check_observer(c);
    return rc ? rc : s->observer_start(c);
}
int fcc_observe(fcc_sim *s) { FCC_HANDLE(s); return s->observe(); }
int fcc_contours(fcc_sim *s, double *X, double *Y) { FCC_HANDLE(s); return s->contours(X, Y); }
int fcc_frame_request(fcc_sim *s) { FCC_HANDLE(s); return s->frame_request(); }
int fcc_frame_read(fcc_sim *s, double *X, double *Y, int64_t *clock) {
    FCC_HANDLE(s);
    return s->frame_read(X, Y, clock);
}
int fcc_checkpoint_save(fcc_sim *s, void *buf, int64_t bytes, int64_t *needed) {
    FCC_HANDLE(s);
    if (bytes < 0) return fail(FCC_ERR_ARGUMENT, "negative size");
    return s->checkpoint(buf, bytes, needed);
}
int fcc_checkpoint_load(fcc_sim *s, const void *buf, int64_t bytes) {
    FCC_HANDLE_MUT(s);
    return s->restore(buf, bytes);
}
int fcc_run_observed(fcc_sim *s, int64_t n, int64_t every) {
    FCC_HANDLE(s);
    if (n < 0 || every < 1) return fail(FCC_ERR_ARGUMENT, "need n >= 0 and every >= 1");
    for (int64_t done = 0; done < n;) {
        const int64_t k = std::min(every, n - done);
        int rc;
        if ((rc = s->run(k)) || (rc = s->observe())) return rc;
        done += k;
    }
    return FCC_OK;
}
int fcc_observer_read(fcc_sim *s, int32_t *ns, int64_t *clock, double *msd, double *pm, double *p2, uint64_t *hist) {
    FCC_HANDLE(s);
    return s->observer_read(ns, clock, msd, pm, p2, hist);
}

int fcc_philox_noise(uint64_t seed, uint64_t draw, int64_t n, int32_t precision, void *out) {
    if (n < 0 || !out) return fail(FCC_ERR_ARGUMENT, "bad noise arguments");
    if (precision != FCC_FP64 && precision != FCC_FP32) return fail(FCC_ERR_ARGUMENT, "precision must be FP64 or FP32");
    if (n == 0) return FCC_OK;
    const size_t bytes = (size_t)3 * n * (precision == FCC_FP64 ? 8 : 4);
    void *dv = nullptr;
    HIPCHK(hipMalloc(&dv, bytes));
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    if (precision == FCC_FP64)
        hipLaunchKernelGGL((k_noise<double>), blocks(n, 256), 256, 0, 0, n, k0, k1, draw, (double *)dv);
    else
        hipLaunchKernelGGL((k_noise<float>), blocks(n, 256), 256, 0, 0, n, k0, k1, draw, (float *)dv);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(out, dv, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(dv);
    HIPCHK(e);
    return FCC_OK;
}

// Staging buffer of fcc_philox_noise_noalloc: a fixed array of the code
// object (3 normals of FCC_NOALLOC_MAX_NOISE bodies in FP64, 6 MiB), one copy
// per GPU, held until the process ends. Its host-side name is not a device
// address: host code uses the address hipGetSymbolAddress returns for the
// current device (noalloc_scratch).
static __device__ __align__(256) double scratch[3 * FCC_NOALLOC_MAX_NOISE];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];

// Device address of `scratch` on the current device, looked up on the first
// call on that device (with it current) and kept per device.
static int noalloc_scratch(void **out) {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(FCC_ERR_DEVICE, "device %d beyond the %d devices of the scratch table", dev, NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(FCC_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return FCC_OK;
}

int fcc_philox_noise_noalloc(uint64_t seed, uint64_t draw, int64_t n, int32_t precision, void *out) {
    if (n < 0 || !out) return fail(FCC_ERR_ARGUMENT, "bad noise arguments");
    if (precision != FCC_FP64 && precision != FCC_FP32) return fail(FCC_ERR_ARGUMENT, "precision must be FP64 or FP32");
    if (n > FCC_NOALLOC_MAX_NOISE)
        return fail(FCC_ERR_ARGUMENT, "n = %lld exceeds the %d bodies of the fixed noise buffer",
                    (long long)n, FCC_NOALLOC_MAX_NOISE);
    if (n == 0) return FCC_OK;
    void *dv = nullptr;
    int rc = noalloc_scratch(&dv);
    if (rc) return rc;
    const size_t bytes = (size_t)3 * n * (precision == FCC_FP64 ? 8 : 4);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    if (precision == FCC_FP64)
        hipLaunchKernelGGL((k_noise<double>), blocks(n, 256), 256, 0, 0, n, k0, k1, draw, (double *)dv);
    else
        hipLaunchKernelGGL((k_noise<float>), blocks(n, 256), 256, 0, 0, n, k0, k1, draw, (float *)dv);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out, dv, bytes, hipMemcpyDeviceToHost));
    return FCC_OK;
}

int fcc_philox4x32(const uint32_t *ctr, const uint32_t *key, int64_t n, uint32_t *out) {
    if (n < 0 || !ctr || !key || !out) return fail(FCC_ERR_ARGUMENT, "bad philox arguments");
    if (n == 0) return FCC_OK;
    uint32_t *dc = nullptr, *dk = nullptr, *dout = nullptr;
    HIPCHK(hipMalloc(&dc, 16 * n));
    HIPCHK(hipMalloc(&dk, 8 * n));
    HIPCHK(hipMalloc(&dout, 16 * n));
    hipError_t e = hipMemcpy(dc, ctr, 16 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dk, key, 8 * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_philox, blocks(n, 256), 256, 0, 0, n, dc, dk, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, 16 * n, hipMemcpyDeviceToHost);
    (void)hipFree(dc);
    (void)hipFree(dk);
    (void)hipFree(dout);
    HIPCHK(e);
    return FCC_OK;
}

// ---- ensembles

int fcc_ensemble_create(int32_t n, const fcc_config *cfgs, const fcc_bodies *bodies, fcc_ensemble **out) {
    if (!out) return fail(FCC_ERR_ARGUMENT, "null output handle");
    *out = nullptr;
    if (n < 1 || !cfgs || !bodies) return fail(FCC_ERR_ARGUMENT, "an ensemble needs at least one replica");
    for (int32_t r = 0; r < n; r++) {
        int rc = validate(&cfgs[r], &bodies[r]);
        if (rc) {
            char msg[sizeof g_err];
            std::snprintf(msg, sizeof msg, "%s", g_err);
            return fail(rc, "replica %d: %s", r, msg);
        }
        const fcc_config &a = cfgs[0], &c = cfgs[r];
        if (c.n_modes != a.n_modes || c.n_theta != a.n_theta || c.precision != a.precision || c.flags != a.flags)
            return fail(FCC_ERR_UNSUPPORTED, "replica %d: n_modes, n_theta, precision and flags must equal replica 0's", r);
    }
    fcc_ensemble *e = nullptr;
    int rc;
    if (cfgs[0].precision == FCC_FP64) {
        auto *q = new Ensemble<double>();
        rc = q->init(n, cfgs, bodies);
        e = q;
    } else {
        auto *q = new Ensemble<float>();
        rc = q->init(n, cfgs, bodies);
        e = q;
    }
    if (rc) {
        delete e;
        return rc;
    }
    *out = e;
    return FCC_OK;
}

int fcc_ensemble_destroy(fcc_ensemble *e) {
    delete e;
    return FCC_OK;
}

#define FCC_ENS(e) \
    if (!(e)) return fail(FCC_ERR_ARGUMENT, "null ensemble handle")

int fcc_ensemble_member(fcc_ensemble *e, int32_t r, fcc_sim **member) {
    FCC_ENS(e);
    if (!member) return fail(FCC_ERR_ARGUMENT, "null output handle");
    if (r < 0 || r >= e->size()) return fail(FCC_ERR_ARGUMENT, "replica %d outside 0:%d", r, e->size() - 1);
    *member = e->member(r);
    return FCC_OK;
}

int fcc_ensemble_run(fcc_ensemble *e, int64_t n) {
    FCC_ENS(e);
    if (n < 0) return fail(FCC_ERR_ARGUMENT, "steps must be non-negative");
    return e->run(n);
}

int fcc_ensemble_run_timed(fcc_ensemble *e, int64_t n, double *fms, double *ums) {
    FCC_ENS(e);
    if (n < 0) return fail(FCC_ERR_ARGUMENT, "steps must be non-negative");
    return e->run_timed(n, fms, ums);
}

int fcc_ensemble_get_launch(fcc_ensemble *e, fcc_launch *lc) { FCC_ENS(e); return e->get_launch(lc); }
int fcc_ensemble_set_launch(fcc_ensemble *e, const fcc_launch *lc) { FCC_ENS(e); return e->set_launch(lc); }

int fcc_ensemble_autotune(fcc_ensemble *e, int64_t steps, fcc_launch *best, double *ms) {
    FCC_ENS(e);
    return e->autotune(steps, best, ms);
}

int fcc_ensemble_observer_start(fcc_ensemble *e, const fcc_observer_config *c) {
    FCC_ENS(e);
    int rc = check_observer(c);
    for (int32_t r = 0; r < e->size() && !rc; r++) rc = e->member(r)->observer_start(c);
    return rc;
}

int fcc_ensemble_observe(fcc_ensemble *e) {
    FCC_ENS(e);
    int rc = FCC_OK;
    for (int32_t r = 0; r < e->size() && !rc; r++) rc = e->member(r)->observe();
    return rc;
}

int fcc_ensemble_run_observed(fcc_ensemble *e, int64_t n, int64_t every) {
    FCC_ENS(e);
    if (n < 0 || every < 1) return fail(FCC_ERR_ARGUMENT, "need n >= 0 and every >= 1");
    for (int64_t done = 0; done < n;) {
        const int64_t k = std::min(every, n - done);
        int rc;
        if ((rc = e->run(k)) || (rc = fcc_ensemble_observe(e))) return rc;
        done += k;
    }
    return FCC_OK;
}

}  // extern "C"
