// fcc_ensemble.hpp — ensembles of independent simulations (host side of
// fcc_ensemble_* in include/fcc.h; included by fcc_api.hip after Sim).
//
// The reference advances independent simulations concurrently, one task and
// one stream each (README.md:249-261, test/test_cuda.jl:95-104); a small
// system of a few thousand bodies leaves most of the GPU idle. Here the
// replicas of an ensemble are ordinary simulations whose buffers stay where
// each Sim allocated them, sharing one stream and one contiguous array of
// status slots. A step of all replicas is one launch per kernel: the grid is
// a table of tiles (replica, workgroup of that replica) and every workgroup
// reads its replica's Params / Ptrs from device tables, so replicas of any
// size share one launch. The step is captured once in a hipGraph (the graph
// reads the tables at run time, so new list capacities or dt need no new
// capture). A replica whose Verlet skin trips is frozen by its own status
// flag until the host rebuilds its list between graph chunks, and a
// per-replica step budget freezes replicas that have finished, so every
// replica executes exactly the kernels fcc_run would give it alone.
//
// Where every replica would run one kernel per step alone (two lanes per
// body, FP32 or mixed without Bending, no parent drag, bonds or debug
// buffers), an ensemble step is k_gate and one step kernel over tiles of 64
// bodies: k_gate also flips the replica's buffer parity (`cur`) for a step
// that executes, and the step kernel reads the other parity's positions, lab
// rows and radius bounds and writes `cur`'s. Between runs every replica's
// current buffers are its first ones and `cur` is 0 (normalize()).

struct fcc_ensemble {
    virtual ~fcc_ensemble() {}
    virtual int32_t size() const = 0;
    virtual fcc_sim *member(int32_t r) = 0;
    virtual int run(int64_t n) = 0;
    virtual int run_timed(int64_t n, double *fms, double *ums) = 0;
    virtual int get_launch(fcc_launch *lc) = 0;
    virtual int set_launch(const fcc_launch *lc) = 0;
    virtual int autotune(int64_t steps, fcc_launch *best, double *ms) = 0;
};

static int check_launch(const fcc_launch *lc) {
    if (!lc) return fail(FCC_ERR_ARGUMENT, "null launch shape");
    if (lc->update_block != 64 && lc->update_block != 128 && lc->update_block != 256)
        return fail(FCC_ERR_ARGUMENT, "update_block must be 64, 128 or 256");
    if (lc->xcd_order != 0 && lc->xcd_order != 1) return fail(FCC_ERR_ARGUMENT, "xcd_order must be 0 or 1");
    if (lc->chunk < 0) return fail(FCC_ERR_ARGUMENT, "chunk must be non-negative");
    if (lc->lanes != 0 && lc->lanes != 2) return fail(FCC_ERR_ARGUMENT, "lanes must be 0 or 2");
    if (lc->kernels != 0 && lc->kernels != 2) return fail(FCC_ERR_ARGUMENT, "kernels must be 0 or 2");
    return FCC_OK;
}

// Tiles of one launch: ceil(N_r / tile) workgroups per replica, in replica
// order. With `xcd` the table is padded to a multiple of 8 and permuted so
// that workgroup b, which the dispatcher places on XCD b mod 8, takes tile
// (b mod 8)·G/8 + b/8: each XCD walks one contiguous range of tiles and a
// replica's neighbour gathers stay within one XCD's L2.
static std::vector<Tile> tile_table(const std::vector<int64_t> &sizes, int tile, bool xcd) {
    std::vector<Tile> lin;
    for (size_t r = 0; r < sizes.size(); r++) {
        const int64_t nb = (sizes[r] + tile - 1) / tile;
        for (int64_t b = 0; b < nb; b++) lin.push_back(Tile{(int32_t)r, (int32_t)b});
    }
    if (!xcd) return lin;
    const size_t G = (lin.size() + 7) / 8 * 8, seg = G / 8;
    std::vector<Tile> out(G);
    for (size_t b = 0; b < G; b++) {
        const size_t tau = (b % 8) * seg + b / 8;
        out[b] = tau < lin.size() ? lin[tau] : Tile{-1, 0};
    }
    return out;
}

template <typename T>
struct Ensemble final : fcc_ensemble {
    std::vector<Sim<T> *> m;
    hipStream_t stream = nullptr;
    DevStatus *st = nullptr;  // one slot per replica
    Params<T> *dP = nullptr;
    Ptrs<T> *dD = nullptr;
    Tile *tf = nullptr, *tu = nullptr;  // tiles of the forces / update launches
    size_t cap_f = 0, cap_u = 0;
    // forces tiles: gf4 of the four-lane kernel (MFMA or FP32 / FP64; replicas below
    // FOUR_LANES_BELOW bodies, as each would run alone), then gf of the
    // other kernel
    unsigned gf4 = 0, gf = 0, gu = 0;
    std::vector<Params<T>> hP, upP;
    std::vector<Ptrs<T>> hD, upD;
    bool uploaded = false;
    fcc_launch lc{256, 1, 0, 0, 0};
    fcc_launch tiles_lc{0, 0, 0, 0, 0};
    int var = -1, tiles_var = -1, ncls = 1;  // forces kernel: 3 wide, 2 MFMA, 1 fast, 0 generic
    bool one = false, tiles_one = false;     // one kernel per step (k_step_*_ens)
    bool bonds = false;                      // a replica has bond terms on bonds: bond kernel after forces
    // captured steps: gx[b] holds 2^b steps (1 to 32), so k steps take
    // k / 32 + popcount(k mod 32) graph launches
    static constexpr int NGE = 6;
    hipGraphExec_t gx[NGE] = {};
    bool graph_ok = false;
    std::vector<double> interval;  // mean steps between rebuilds, per replica
    std::vector<int64_t> since;
    bool want_timed = false;
    double acc_f = 0.0, acc_u = 0.0;
    int64_t acc_n = 0;
    // pipelined runs: list builds on a side stream, status snapshots of the
    // chunks in flight in pinned host memory
    static constexpr int NB = 4;
    hipStream_t side = nullptr;
    DevStatus *snap = nullptr;
    hipEvent_t sev[NB] = {};
    hipEvent_t tev[NB][4] = {};
    bool sone[NB] = {};  // the chunk in the slot was issued as one-kernel steps (no update events)

    ~Ensemble() override {
        if (stream) (void)hipStreamSynchronize(stream);
        if (side) (void)hipStreamSynchronize(side);
        for (hipGraphExec_t g : gx)
            if (g) (void)hipGraphExecDestroy(g);
        for (int q = 0; q < NB; q++) {
            if (sev[q]) (void)hipEventDestroy(sev[q]);
            for (hipEvent_t e : tev[q])
                if (e) (void)hipEventDestroy(e);
        }
        if (snap) (void)hipHostFree(snap);
        for (Sim<T> *s : m) delete s;
        for (void *q : {(void *)st, (void *)dP, (void *)dD, (void *)tf, (void *)tu})
            if (q) (void)hipFree(q);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int32_t size() const override { return (int32_t)m.size(); }
    fcc_sim *member(int32_t r) override { return m[r]; }

    int init(int32_t n, const fcc_config *cfgs, const fcc_bodies *bodies) {
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIPCHK(hipMalloc(&st, sizeof(DevStatus) * n));
        HIPCHK(hipMalloc(&dP, sizeof(Params<T>) * n));
        HIPCHK(hipMalloc(&dD, sizeof(Ptrs<T>) * n));
        for (int32_t r = 0; r < n; r++) {
            auto *s = new Sim<T>();
            s->borrowed = true;
            m.push_back(s);
            int rc = s->init(&cfgs[r], &bodies[r], stream, st + r);
            if (rc) return rc;
        }
        const Sim<T> &a = *m[0];
        for (int32_t r = 0; r < n; r++) {
            const Sim<T> &s = *m[r];
            if (s.has_pdrag != a.has_pdrag)
                return fail(FCC_ERR_UNSUPPORTED, "replica %d: parent drag must be present in all replicas or none", r);
            if (s.mfma_T != a.mfma_T)
                return fail(FCC_ERR_UNSUPPORTED, "replica %d: Bending must be present in all replicas or none", r);
            if (s.has_contact && s.skin <= 0.0)
                return fail(FCC_ERR_UNSUPPORTED, "replica %d: contact without a Verlet skin has no graph step", r);
        }
        hP.resize(n);
        hD.resize(n);
        interval.assign(n, 64.0);
        since.assign(n, 0);
        HIPCHK(hipStreamSynchronize(stream));
        return FCC_OK;
    }

    // ------------------------------------------------------------ launches

    int launch_gate() {
        const int R = (int)m.size();
        hipLaunchKernelGGL(k_gate, blocks(R, 64), 64, 0, stream, st, R, one ? 1 : 0);
        HIPCHK(hipGetLastError());
        return FCC_OK;
    }

    // Whether every replica would run one kernel per step alone (var: the
    // forces kernel variant of prepare()).
    bool one_kernel(int v) const {
        if constexpr (sizeof(T) != 4) return false;
        const Sim<T> &a = *m[0];
        if ((v != 1 && v != 2) || lc.kernels == 2) return false;
        if (a.nth != 40 || a.has_pdrag || a.debug || (a.cfg.flags & FCC_FLAG_TWO_KERNELS)) return false;
        if (v == 2 && a.mfma_T != 3) return false;
        for (const Sim<T> *s : m)
            if (!s->has_contact || s->skin <= 0.0 || s->has_bonds) return false;
        return true;
    }

    // The step kernels of a one-kernel ensemble step (tiles of 64 bodies): the
    // four-lane one over the gf4 tiles of the replicas below four_below
    // bodies, the two-lane one over the gf tiles after them.
    int launch_step_kernel(hipEvent_t eb = nullptr, hipEvent_t ee = nullptr) {
        if constexpr (sizeof(T) == 4) {
            const Sim<T> &a = *m[0];
            hipEvent_t e4 = gf ? nullptr : ee, e2 = gf4 ? nullptr : eb;
            if (var == 2) {
                const bool bf = a.cfg.precision == FCC_MIXED_BF16;
#define FCC_SME2(KV, NCV, BFV)                                                                                \
    if (gf4) FCC_LAUNCH((k_step_mfma4_ens<KV, 3, NCV, BFV>), gf4, 256, 0, stream, eb, e4, dP, dD, tf, a.asyn4, a.bpull4); \
    if (gf) FCC_LAUNCH((k_step_mfma_ens<KV, 3, NCV, BFV>), gf, 128, 0, stream, e2, ee, dP, dD, tf + gf4, a.asyn, a.bpull);
#define FCC_SME(KV, NCV) if (bf) { FCC_SME2(KV, NCV, true) } else { FCC_SME2(KV, NCV, false) }
                if (a.K == 5) {
                    if (ncls <= 1) { FCC_SME(5, 1) } else { FCC_SME(5, 2) }
                } else {
                    if (ncls <= 1) { FCC_SME(7, 1) } else { FCC_SME(7, 2) }
                }
#undef FCC_SME
#undef FCC_SME2
            } else {
#define FCC_SFE(KV, NCV)                                                                                      \
    if (gf4) FCC_LAUNCH((k_step_fast4_ens<T, KV, NCV>), gf4, 256, 0, stream, eb, e4, dP, dD, tf);             \
    if (gf) FCC_LAUNCH((k_step_fast2_ens<T, KV, NCV>), gf, 128, 0, stream, e2, ee, dP, dD, tf + gf4);
                if (a.K == 5) {
                    if (ncls <= 1) { FCC_SFE(5, 1) } else { FCC_SFE(5, 2) }
                } else {
                    if (ncls <= 1) { FCC_SFE(7, 1) } else { FCC_SFE(7, 2) }
                }
#undef FCC_SFE
            }
            HIPCHK(hipGetLastError());
            return FCC_OK;
        }
        return fail(FCC_ERR_STATE, "one-kernel ensemble step in double precision");
    }

    // Every replica's current buffers named as its first ones, parity 0 on
    // the device, and the tables uploaded to match (the device is idle).
    int normalize() {
        if (!one && !tiles_one) return FCC_OK;
        std::vector<DevStatus> hs;
        int rc = read_status(hs);
        if (rc) return rc;
        bool any = false;
        for (size_t r = 0; r < m.size(); r++)
            if (hs[r].cur != 0) {
                m[r]->flip();
                any = true;
            }
        if (!any) return FCC_OK;
        const int R = (int)m.size();
        hipLaunchKernelGGL(k_clear_parity, blocks(R, 64), 64, 0, stream, st, R);
        HIPCHK(hipGetLastError());
        return prepare();
    }

    // Forces of one ensemble step, then the bond kernel over the update tiles
    // when a replica has bonds (replicas without bonds return at once).
    int launch_forces(hipEvent_t eb = nullptr, hipEvent_t ee = nullptr) {
        int rc = launch_forces_kernel(eb, ee);
        if (rc || !bonds) return rc;
        hipLaunchKernelGGL((k_bonds_ens<T>), gu, lc.update_block, 0, stream, dP, dD, tu, 2);
        HIPCHK(hipGetLastError());
        return FCC_OK;
    }

    // The MFMA and fast kernels count hits per node in NPLANE bit planes.
    bool fast_var() const { return var == 1 || var == 2; }

    int launch_forces_kernel(hipEvent_t eb, hipEvent_t ee) {
        const Sim<T> &a = *m[0];
        const int dbg = a.debug ? 1 : 0;
        if constexpr (sizeof(T) == 4) {
            if (var == 2) {
                const bool bf = a.cfg.precision == FCC_MIXED_BF16;
#define FCC_MFE2(KV, TV, NCV, BFV)                                                                            \
    if (gf4) FCC_LAUNCH((k_forces_mfma4_ens<KV, TV, NCV, BFV>), gf4, 64, 0, stream, eb, gf ? nullptr : ee, dP, dD, \
                        tf, a.asyn4, a.bpull4, 2, dbg);                                                       \
    if (gf) FCC_LAUNCH((k_forces_mfma_ens<KV, TV, NCV, BFV>), gf, 64, 0, stream, gf4 ? nullptr : eb, ee, dP, dD, \
                       tf + gf4, a.asyn, a.bpull, 2, dbg);
#define FCC_MFE(KV, TV, NCV) if (bf) { FCC_MFE2(KV, TV, NCV, true) } else { FCC_MFE2(KV, TV, NCV, false) }
#define FCC_MFE_T(KV, NCV) if (a.mfma_T == 3) { FCC_MFE(KV, 3, NCV) } else { FCC_MFE(KV, 4, NCV) }
                if (a.K == 5) {
                    if (ncls <= 1) { FCC_MFE_T(5, 1) } else { FCC_MFE_T(5, 2) }
                } else {
                    if (ncls <= 1) { FCC_MFE_T(7, 1) } else { FCC_MFE_T(7, 2) }
                }
#undef FCC_MFE_T
#undef FCC_MFE
#undef FCC_MFE2
                HIPCHK(hipGetLastError());
                return FCC_OK;
            }
        }
        if (var == 1) {
#define FCC_FE(KV, NC)                                                                                        \
    if (a.nth == 40) {                                                                                        \
        if (gf4) FCC_LAUNCH((k_forces_fast2_ens<T, KV, NC, 4>), gf4, 64, 0, stream, eb, gf ? nullptr : ee, dP, dD, \
                            tf, 2, dbg);                                                                      \
        if (gf) FCC_LAUNCH((k_forces_fast2_ens<T, KV, NC, 2>), gf, 64, 0, stream, gf4 ? nullptr : eb, ee, dP, dD, \
                           tf + gf4, 2, dbg);                                                                 \
    } else                                                                                                    \
        FCC_LAUNCH((k_forces_fast_ens<T, 0, KV, NC>), gf, 64, 0, stream, eb, ee, dP, dD, tf, 2, dbg);
            if (a.K == 5) {
                if (ncls <= 1) { FCC_FE(5, 1) } else { FCC_FE(5, 2) }
            } else {
                if (ncls <= 1) { FCC_FE(7, 1) } else { FCC_FE(7, 2) }
            }
#undef FCC_FE
            HIPCHK(hipGetLastError());
            return FCC_OK;
        }
        if (var == 3) {
            const size_t sh = sizeof(T) * (size_t)a.nth;
#define FCC_WE(KMV)                                                                                           \
    if (a.nth <= WIDE_WG64_THETA)                                                                             \
        FCC_LAUNCH((k_forces_wide_ens<T, KMV, 64>), gf, 64, sh, stream, eb, ee, dP, dD, tf, 2, dbg);            \
    else FCC_LAUNCH((k_forces_wide_ens<T, KMV, 256>), gf, 256, sh, stream, eb, ee, dP, dD, tf, 2, dbg);
            switch (a.KM) {
            case 5: FCC_WE(5) break;
            case 7: FCC_WE(7) break;
            case 15: FCC_WE(15) break;
            default: FCC_WE(31)
            }
#undef FCC_WE
            HIPCHK(hipGetLastError());
            return FCC_OK;
        }
#define FCC_GE(KMV)                                                                                           \
    if (a.nth == 40) FCC_LAUNCH((k_forces_ens<T, 40, KMV>), gf, 64, 0, stream, eb, ee, dP, dD, tf, 2, dbg);   \
    else FCC_LAUNCH((k_forces_ens<T, 0, KMV>), gf, 64, 0, stream, eb, ee, dP, dD, tf, 2, dbg);
        switch (a.KM) {
        case 5: FCC_GE(5) break;
        case 7: FCC_GE(7) break;
        case 15: FCC_GE(15) break;
        default: FCC_GE(31)
        }
#undef FCC_GE
        HIPCHK(hipGetLastError());
        return FCC_OK;
    }

    int launch_update(hipEvent_t eb = nullptr, hipEvent_t ee = nullptr) {
        const Sim<T> &a = *m[0];
        const int pd = a.has_pdrag ? 1 : 0, dbg = a.debug ? 1 : 0;
        const int ub = lc.update_block;
#define FCC_UE(KMV, EX)                                                                                       \
    if (a.mixed) FCC_LAUNCH((k_update_ens<T, KMV, EX, true>), gu, ub, 0, stream, eb, ee, dP, dD, tu, pd, dbg); \
    else FCC_LAUNCH((k_update_ens<T, KMV, EX, false>), gu, ub, 0, stream, eb, ee, dP, dD, tu, pd, dbg);
        if (a.K == 5) {
            FCC_UE(5, true)
        } else if (a.K == 7) {
            FCC_UE(7, true)
        } else if (a.KM == 5) {
            FCC_UE(5, false)
        } else if (a.KM == 7) {
            FCC_UE(7, false)
        } else if (a.KM == 15) {
            FCC_UE(15, false)
        } else {
            FCC_UE(31, false)
        }
#undef FCC_UE
        HIPCHK(hipGetLastError());
        return FCC_OK;
    }

    // Tables, tiles and the captured step, brought up to date with the
    // replicas (kernel variant, list capacities, dt) and the launch shape.
    int prepare() {
        bool fast = true, bd = false;
        int nc = 1;
        for (Sim<T> *s : m) {
            fast = fast && s->fast_ok();
            nc = std::max(nc, s->p.ncls);
            bd = bd || s->has_bonds;
        }
        const int v = fast ? ((sizeof(T) == 4 && m[0]->mfma_T) ? 2 : 1) : (m[0]->nth > FUSED_MAX_THETA ? 3 : 0);
        if (v != var || nc != ncls || bd != bonds) graph_ok = false;
        var = v;
        ncls = nc;
        bonds = bd;
        const size_t R = m.size();
        const bool o = one_kernel(v);
        if (o != one) graph_ok = false;
        one = o;
        if (one)
            for (Sim<T> *s : m)
                if (int rc = s->ensure_next()) return rc;
        for (size_t r = 0; r < R; r++) {
            Sim<T> &s = *m[r];
            s.p.skin_check = (s.has_contact && s.skin > 0.0) ? 1 : 0;
            hP[r] = s.p;
            hD[r] = s.d;
        }
        if (!uploaded || std::memcmp(hP.data(), upP.data(), sizeof(Params<T>) * R) != 0 ||
            std::memcmp(hD.data(), upD.data(), sizeof(Ptrs<T>) * R) != 0) {
            upP = hP;
            upD = hD;
            HIPCHK(hipMemcpyAsync(dP, upP.data(), sizeof(Params<T>) * R, hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(dD, upD.data(), sizeof(Ptrs<T>) * R, hipMemcpyHostToDevice, stream));
            uploaded = true;
        }
        if (var != tiles_var || one != tiles_one || lc.update_block != tiles_lc.update_block ||
            lc.xcd_order != tiles_lc.xcd_order || lc.lanes != tiles_lc.lanes) {
            std::vector<int64_t> sizes(R), small(R, 0), large(R, 0);
            for (size_t r = 0; r < R; r++) sizes[r] = m[r]->N;
            // Bodies per forces workgroup: 16 for the four-lane kernels (replicas
            // below FOUR_LANES_BELOW bodies at nθ = 40, the kernel each would get
            // alone: MFMA or FP32 / FP64), 32 for the two-lane kernels, one for
            // contours of more than 64 nodes, 64 otherwise; 64 for both lane
            // counts in one-kernel steps.
            const bool fast40 = (var == 1 && m[0]->nth == 40) || var == 2;
            for (size_t r = 0; r < R; r++) (fast40 && sizes[r] < m[r]->four_below ? small : large)[r] = sizes[r];
            const int tile = one ? 64 : (var == 3 ? 1 : (var == 2 || fast40 ? 32 : 64));
            std::vector<Tile> a = tile_table(small, one ? 64 : 16, lc.xcd_order != 0);
            const size_t n4 = a.size();
            const std::vector<Tile> a2 = tile_table(large, tile, lc.xcd_order != 0);
            a.insert(a.end(), a2.begin(), a2.end());
            const std::vector<Tile> b = tile_table(sizes, lc.update_block, lc.xcd_order != 0);
            HIPCHK(hipStreamSynchronize(stream));
            if (a.size() > cap_f) {
                if (tf) HIPCHK(hipFree(tf));
                tf = nullptr;
                HIPCHK(hipMalloc(&tf, sizeof(Tile) * a.size()));
                cap_f = a.size();
            }
            if (b.size() > cap_u) {
                if (tu) HIPCHK(hipFree(tu));
                tu = nullptr;
                HIPCHK(hipMalloc(&tu, sizeof(Tile) * b.size()));
                cap_u = b.size();
            }
            HIPCHK(hipMemcpy(tf, a.data(), sizeof(Tile) * a.size(), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(tu, b.data(), sizeof(Tile) * b.size(), hipMemcpyHostToDevice));
            gf4 = (unsigned)n4;
            gf = (unsigned)(a.size() - n4);
            gu = (unsigned)b.size();
            tiles_var = var;
            tiles_one = one;
            tiles_lc = lc;
            graph_ok = false;
        }
        if (!graph_ok) {
            std::lock_guard<std::mutex> lk(g_capture_lock);
            for (hipGraphExec_t &g : gx)
                if (g) {
                    HIPCHK(hipGraphExecDestroy(g));
                    g = nullptr;
                }
            int rc = FCC_OK;
            for (int b = 0; b < NGE && !rc; b++) rc = capture_steps(1 << b, &gx[b]);
            if (rc) return rc;
            graph_ok = true;
        }
        return FCC_OK;
    }

    int capture_steps(int steps, hipGraphExec_t *out) {
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = FCC_OK;
        for (int q = 0; q < steps && !rc; q++) {
            rc = launch_gate();
            if (one) {
                if (!rc) rc = launch_step_kernel();
                continue;
            }
            if (!rc) rc = launch_forces();
            if (!rc) rc = launch_update();
        }
        hipError_t e = hipStreamEndCapture(stream, &graph);
        if (rc) return rc;
        HIPCHK(e);
        HIPCHK(hipGraphInstantiate(out, graph, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(graph));
        HIPCHK(hipGraphUpload(*out, stream));  // its first replay in a run starts at once
        return FCC_OK;
    }

    // k graph steps: k / 32 launches of the longest graph, then one graph per
    // set bit of k mod 32.
    int replay(int64_t k) {
        constexpr int64_t LONG = 1 << (NGE - 1);
        for (int64_t q = 0; q < k / LONG; q++) HIPCHK(hipGraphLaunch(gx[NGE - 1], stream));
        for (int b = NGE - 2; b >= 0; b--)
            if ((k % LONG) & (1 << b)) HIPCHK(hipGraphLaunch(gx[b], stream));
        return FCC_OK;
    }

    int read_status(std::vector<DevStatus> &hs) {
        hs.resize(m.size());
        HIPCHK(hipMemcpyAsync(hs.data(), st, sizeof(DevStatus) * m.size(), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return FCC_OK;
    }

    int set_budgets(int64_t steps) {
        const int R = (int)m.size();
        hipLaunchKernelGGL(k_ensemble_targets, blocks(R, 64), 64, 0, stream, st, R, steps);
        HIPCHK(hipGetLastError());
        return FCC_OK;
    }

    // update_neighbors! of every replica whose list is stale or whose skin
    // tripped on its last step (run_steps! begins with it).
    int refresh_lists(std::vector<DevStatus> &hs) {
        int rc = read_status(hs);
        if (rc) return rc;
        for (size_t r = 0; r < m.size(); r++) {
            Sim<T> &s = *m[r];
            if (s.has_contact && (!s.valid || hs[r].rebuild != 0) && (rc = s.rebuild())) return rc;
        }
        return FCC_OK;
    }

    // Steps per chunk: an eighth of the shortest rebuild-interval estimate
    // (a tripped replica stays frozen until the chunk after its trip has
    // been read), 8 to 32 steps, rounded down to a power of two (one graph
    // launch per chunk).
    int64_t chunk_len() const {
        if (lc.chunk > 0) return lc.chunk;
        double mi = 1e300;
        for (size_t r = 0; r < m.size(); r++)
            if (m[r]->has_contact) mi = std::min(mi, interval[r]);
        if (mi >= 1e300) return 64;
        const int64_t c = std::min<int64_t>(4 * GRAPH_STEPS, std::max<int64_t>(GRAPH_STEPS, (int64_t)std::llround(mi / 8)));
        int64_t p2 = GRAPH_STEPS;
        while (2 * p2 <= c) p2 *= 2;
        return p2;
    }

    int ensure_async() {
        if (side) return FCC_OK;
        HIPCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        HIPCHK(hipHostMalloc((void **)&snap, sizeof(DevStatus) * NB * m.size(), hipHostMallocDefault));
        for (int q = 0; q < NB; q++) {
            HIPCHK(hipEventCreateWithFlags(&sev[q], hipEventDisableTiming));
            for (auto &e : tev[q]) HIPCHK(hipEventCreate(&e));
        }
        return FCC_OK;
    }

    // k steps on the main stream, then a snapshot of every status slot.
    int issue(int64_t k, int slot) {
        const size_t R = m.size();
        if (!want_timed) {
            int rc = replay(k);
            if (rc) return rc;
        } else {
            // the chunk's last step eagerly between events: the kernels and
            // arguments of the captured step
            int rc0 = replay(k - 1);
            if (rc0) return rc0;
            int rc = launch_gate();
            if (rc) return rc;
            sone[slot] = one;
            if (one) {
                if ((rc = launch_step_kernel(tev[slot][0], tev[slot][1]))) return rc;
            } else {
                if ((rc = launch_forces(tev[slot][0], tev[slot][1]))) return rc;
                if ((rc = launch_update(tev[slot][2], tev[slot][3]))) return rc;
            }
        }
        HIPCHK(hipMemcpyAsync(snap + (size_t)slot * R, st, sizeof(DevStatus) * R, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipEventRecord(sev[slot], stream));
        return FCC_OK;
    }

    // run! of every replica, pipelined: the host keeps two chunks queued on
    // the main stream and reads each chunk's status snapshot while the next
    // one runs. A replica whose skin tripped stays frozen (its workgroups
    // return at once) while its list is rebuilt on the side stream; the gate
    // kernel of the first step after the build's release lets it go on, so
    // it executes the same steps and builds as alone, and the device never
    // waits for the host. Capacity overflows and kernel switches (rare) are
    // handled synchronously.
    int advance(int64_t n) {
        std::vector<DevStatus> hs;
        int rc = refresh_lists(hs);
        if (rc) return rc;
        if ((rc = set_budgets(n)) || (rc = read_status(hs)) || (rc = ensure_async()) || (rc = prepare())) return rc;
        const size_t R = m.size();
        std::vector<int64_t> clock(R), target(R);
        std::vector<int64_t> trip(R);  // clock at the last build issued
        std::vector<int32_t> seen(R), issued_builds(R);
        for (size_t r = 0; r < R; r++) {
            clock[r] = trip[r] = hs[r].clock;
            target[r] = hs[r].target;
            seen[r] = issued_builds[r] = hs[r].builds;
        }
        constexpr int32_t NLIM = 1 << NPLANE;  // neighbour limit of the fast kernels
        int64_t issued = 0, read = 0, stale_before = 0;
        bool finished = false;
        while (!finished || read < issued) {
            while (!finished && issued - read < 2) {
                if ((rc = issue(chunk_len(), (int)(issued % NB)))) return rc;
                issued++;
            }
            const int slot = (int)(read % NB);
            HIPCHK(hipEventSynchronize(sev[slot]));
            const DevStatus *s = snap + (size_t)slot * R;
            const bool stale = read < stale_before;
            read++;
            if (want_timed) {
                float a = 0.f, b = 0.f;
                HIPCHK(hipEventElapsedTime(&a, tev[slot][0], tev[slot][1]));
                if (!sone[slot]) HIPCHK(hipEventElapsedTime(&b, tev[slot][2], tev[slot][3]));
                acc_f += a;
                acc_u += b;
                acc_n++;
            }
            bool all = true;
            std::vector<size_t> special;
            for (size_t r = 0; r < R; r++) {
                Sim<T> &sim = *m[r];
                const int64_t ex = s[r].clock - clock[r];
                clock[r] = s[r].clock;
                for (int64_t q = 0; q < ex; q++) sim.t += sim.cfg.dt;
                sim.nstep += ex;
                since[r] += ex;
                if (s[r].builds != seen[r]) {  // builds completed on the side stream
                    sim.builds += s[r].builds - seen[r];
                    seen[r] = s[r].builds;
                    sim.maxcount = s[r].overflow;
                }
                if (clock[r] < target[r]) all = false;
                if (stale || seen[r] != issued_builds[r] || s[r].rebuild == 0 || clock[r] >= target[r]) continue;
                if (clock[r] == trip[r]) {
                    // no step since the last build: either that build kept the
                    // replica frozen (it did not fit) or its release is not in
                    // this snapshot yet
                    if (s[r].overflow > sim.d.cap || (fast_var() && s[r].overflow >= NLIM)) special.push_back(r);
                    continue;
                }
                // the replica's current buffers: parity s[r].cur (one-kernel steps)
                if (s[r].cur) sim.flip();
                rc = sim.rebuild_async(side, fast_var() ? NLIM : INT32_MAX);
                if (s[r].cur) sim.flip();
                if (rc) return rc;
                issued_builds[r]++;
                trip[r] = clock[r];
                interval[r] = (interval[r] + (double)since[r]) / 2;
                since[r] = 0;
            }
            if (!special.empty()) {
                HIPCHK(hipStreamSynchronize(stream));
                HIPCHK(hipStreamSynchronize(side));
                if ((rc = normalize())) return rc;  // the host's rebuilds use the members' current buffers
                for (size_t r : special) {
                    Sim<T> &sim = *m[r];
                    if (s[r].overflow > sim.d.cap) {
                        sim.builds--;  // the overflowed build is redone, as the regrow loop of Sim::rebuild
                        if ((rc = sim.rebuild())) return rc;
                    } else {
                        sim.maxcount = s[r].overflow;
                        hipLaunchKernelGGL(k_clear_status, 1, 1, 0, stream, sim.d.st);
                        HIPCHK(hipGetLastError());
                    }
                }
                if ((rc = prepare())) return rc;
                stale_before = issued;  // snapshots already queued predate these fixes
            }
            finished = all;
        }
        HIPCHK(hipStreamSynchronize(side));
        return FCC_OK;
    }

    int run(int64_t n) override {
        if (n <= 0) return FCC_OK;
        const int rc = advance(n);
        const int rc2 = set_budgets(-1);  // no budget outside ensemble runs
        hipError_t e = hipStreamSynchronize(stream);
        if (rc) return rc;
        if (rc2) return rc2;
        HIPCHK(e);
        return normalize();
    }

    int run_timed(int64_t n, double *fms, double *ums) override {
        want_timed = true;
        acc_f = acc_u = 0.0;
        acc_n = 0;
        if (int rc = ensure_async()) {
            want_timed = false;
            return rc;
        }
        const int rc = run(n);
        want_timed = false;
        if (rc) return rc;
        if (fms) *fms = acc_n ? acc_f / acc_n : 0.0;
        if (ums) *ums = acc_n ? acc_u / acc_n : 0.0;
        return FCC_OK;
    }

    int get_launch(fcc_launch *out) override {
        if (!out) return fail(FCC_ERR_ARGUMENT, "null launch shape");
        *out = lc;
        return FCC_OK;
    }

    int set_launch(const fcc_launch *in) override {
        int rc = check_launch(in);
        if (rc) return rc;
        lc = *in;
        apply_lanes();
        return FCC_OK;
    }

    // The lane choice of the launch shape on every replica: the four-lane
    // threshold each uses alone and inside the ensemble (a member's own
    // captured step is recaptured when it changes).
    void apply_lanes() {
        for (Sim<T> *s : m) {
            const int64_t fb = (lc.lanes == 2 || (s->cfg.flags & FCC_FLAG_TWO_LANES)) ? 0 : FOUR_LANES_BELOW;
            if (fb != s->four_below) {
                s->four_below = fb;
                s->gen++;
            }
        }
    }

    // Whether the lane choice changes any replica's forces kernel.
    bool lanes_matter() const {
        bool fast = m[0]->nth == 40;
        for (const Sim<T> *s : m) fast = fast && s->fast_ok();
        if (!fast || (m[0]->cfg.flags & FCC_FLAG_TWO_LANES)) return false;
        for (const Sim<T> *s : m)
            if (s->N < FOUR_LANES_BELOW) return true;
        return false;
    }

    // ----------------------------------------------------------- autotuner

    // Device regions a step writes (state, geometry, gradients, status).
    std::vector<std::pair<void *, size_t>> step_regions() {
        std::vector<std::pair<void *, size_t>> v;
        for (Sim<T> *s : m) {
            const size_t N = (size_t)s->N, rows = N * (size_t)s->ldc;
            Ptrs<T> &d = s->d;
            v.push_back({d.x, 8 * N});
            v.push_back({d.y, 8 * N});
            v.push_back({d.imgx, 4 * N});
            v.push_back({d.imgy, 4 * N});
            v.push_back({d.th, sizeof(T) * N});
            v.push_back({d.coef, sizeof(T) * rows});
            v.push_back({d.lab, sizeof(T) * rows});
            v.push_back({d.rmax, sizeof(T) * N});
            v.push_back({d.rmin, sizeof(T) * N});
            v.push_back({d.glab, sizeof(T) * rows});
            v.push_back({d.gx, sizeof(T) * N});
            v.push_back({d.gy, sizeof(T) * N});
            v.push_back({d.gth, sizeof(T) * N});
            v.push_back({d.Ebody, sizeof(T) * N});
        }
        v.push_back({st, sizeof(DevStatus) * m.size()});
        return v;
    }

    int copy_regions(const std::vector<std::pair<void *, size_t>> &v, char *buf, bool save) {
        size_t off = 0;
        for (const auto &q : v) {
            if (save)
                HIPCHK(hipMemcpyAsync(buf + off, q.first, q.second, hipMemcpyDeviceToDevice, stream));
            else
                HIPCHK(hipMemcpyAsync(q.first, buf + off, q.second, hipMemcpyDeviceToDevice, stream));
            off += (q.second + 255) / 256 * 256;
        }
        return FCC_OK;
    }

    int autotune(int64_t steps, fcc_launch *best, double *ms) override {
        if (steps <= 0) return fail(FCC_ERR_ARGUMENT, "steps must be positive");
        std::vector<DevStatus> hs;
        int rc = refresh_lists(hs);
        if (rc) return rc;
        const auto regions = step_regions();
        size_t bytes = 0;
        for (const auto &q : regions) bytes += (q.second + 255) / 256 * 256;
        char *snap = nullptr;
        HIPCHK(hipMalloc(&snap, bytes));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        fcc_launch keep = lc, win = lc;
        double best_ms = 1e300;
        rc = copy_regions(regions, snap, true);
        if (!rc && hipEventCreate(&e0) != hipSuccess) rc = fail(FCC_ERR_DEVICE, "hipEventCreate failed");
        if (!rc && hipEventCreate(&e1) != hipSuccess) rc = fail(FCC_ERR_DEVICE, "hipEventCreate failed");
        const int ubs[3] = {64, 128, 256};
        const int nl = lanes_matter() ? 2 : 1;
        for (int li = 0; li < nl && !rc; li++)
        for (int ki = 0; ki < 2 && !rc; ki++)
        for (int xo = 0; xo < 2 && !rc; xo++) {
            for (int u = 0; u < 3 && !rc; u++) {
                lc = fcc_launch{ubs[u], xo, keep.chunk, nl == 1 ? keep.lanes : 2 * li, 2 * ki};
                apply_lanes();
                if ((rc = prepare())) break;
                // the same candidate as another: one kernel per step (no
                // update workgroup size), or kernels = 0 where it does not apply
                if ((one && u > 0) || (ki == 0 && !one)) continue;
                if ((rc = copy_regions(regions, snap, false))) break;
                hipError_t e = hipSuccess;
                for (int q = 0; q < 2 && e == hipSuccess; q++) e = hipGraphLaunch(gx[0], stream);
                if (e == hipSuccess) e = hipEventRecord(e0, stream);
                if (e == hipSuccess && (rc = replay(steps))) break;
                if (e == hipSuccess) e = hipEventRecord(e1, stream);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float t = 0.f;
                if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
                if (e != hipSuccess) {
                    rc = fail(FCC_ERR_DEVICE, "autotune trial: %s", hipGetErrorString(e));
                    break;
                }
                const double per = (double)t / (double)steps;
                if (per < best_ms) {
                    best_ms = per;
                    win = lc;
                }
            }
        }
        if (!rc) rc = copy_regions(regions, snap, false);
        lc = rc ? keep : win;
        apply_lanes();
        if (!rc) rc = prepare();
        (void)hipStreamSynchronize(stream);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipFree(snap);
        if (rc) return rc;
        if (best) *best = lc;
        if (ms) *ms = best_ms;
        return FCC_OK;
    }
};
