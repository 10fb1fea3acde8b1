// Fused RK4 hot path (turbnstwod.f90:95-168) for one GPU.
//
// Per iteration, six kernels run back to back on the library's stream (captured
// once into a HIP graph and replayed):
//   k_pre        cells: primitives of Q^n, local time step (or the global
//                minimum through an order-independent atomic min), JST
//                pressure / nu~ sensors and undivided Laplacians D2Q
//                (solver_state.f90:77-118, turbnstwod.f90:97-114);
//                nodes: nodal averages of u, v, T, rho, nu~ as a gather in the
//                reference's accumulation order (viscous.f90:40-124).
//   k_dv         JST dissipation D (solver_state.f90:120-223) and viscous
//                fluxes VIS (viscous.f90:62-217); both stay frozen for the four
//                stages, as in the reference.
//   k_stage<s>   s = 1..4: primitives, ghost cells (boundary.f90) computed
//                inline from the parent, source terms, central flux residual
//                and the RK update (residual.f90, source_terms.f90).  Stage 4
//                also writes per-block L2 / Linf partials of C - COLD, which
//                block 0 of the next pre-pass (or k_finalize) folds in a fixed
//                order into the per-iteration history.
// Every kernel is a cell- or node-centric gather: no atomics on values, so
// results are deterministic.  Ghost neighbours of a cell are only ever
// consumed by that cell (its parent), which is what lets the stage kernels
// compute them on the fly instead of running a separate boundary pass.  The
// exception is BC type 6, whose ghost also depends on the mass flow of a whole
// column of cells: on grids that have such ghosts, k_bc6 forms them once per
// stage ahead of the kernels that read them (the kernels are then instantiated
// with that load path; grids without type 6 run kernels compiled without it).
#include "fused_kernels.hpp"

namespace tdt {

hipError_t enqueue_fp64(tdt_ctx *c, FusedBufs &B) {
    if (B.F.ncases) return enqueue_iteration<double, double, double, false, true>(c, B, c->K);
    return B.F.has_bc6 ? enqueue_iteration<double, double, double, true, false>(c, B, c->K)
                       : enqueue_iteration<double, double, double, false, false>(c, B, c->K);
}

namespace {

__global__ void __launch_bounds__(kBlock) k_finalize(FDev F, int nblk) { fold_monitor(F, nblk); }

// Batch: the last iteration's partials, one block per case, then the counter.
__global__ void __launch_bounds__(kBlock) k_finalize_batch(FDev F) { fold_case(F, blockIdx.x); }

__global__ void k_batch_tick(int *ctr) {
    if (ctr[3]) {
        ctr[0] += 1;
        ctr[3] = 0;
    }
}

__global__ void k_reset_counters(int *ctr, unsigned long long *dtslot) {
    ctr[0] = 0;
    ctr[1] = 0;
    ctr[2] = 0;
    ctr[3] = 0;
    *dtslot = (unsigned long long)__double_as_longlong(100000.0);
}

__global__ void __launch_bounds__(kBlock) k_narrow(const double *src, float *dst, int count) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < count) dst[i] = (float)src[i];
}

// node coordinates as pairs (nxy), in the storage type the viscous faces read
template <class T>
__global__ void __launch_bounds__(kBlock) k_node_xy(const void *x, const void *y, void *xy,
                                                     int m) {
    int v = blockIdx.x * kBlock + threadIdx.x;
    if (v < m) {
        static_cast<T *>(xy)[2 * v] = static_cast<const T *>(x)[v];
        static_cast<T *>(xy)[2 * v + 1] = static_cast<const T *>(y)[v];
    }
}

__global__ void __launch_bounds__(kBlock) k_cold_from_qn(FDev F, double *cold) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < F.n)
        for (int q = 0; q < 5; ++q) cold[q * F.n + i] = F.qn[q * F.nt + i];
}

hipError_t enqueue(tdt_ctx *c, FusedBufs &B) {
    return c->precision == TDT_FP64 ? enqueue_fp64(c, B) : enqueue_mixed(c, B);
}

template <class T>
T *dalloc(FusedBufs &B, size_t count) {
    void *p = nullptr;
    if (hipMalloc(&p, count * sizeof(T) + 256) != hipSuccess) return nullptr;
    B.own.push_back(p);
    return static_cast<T *>(p);
}

// Frees one buffer dalloc made (the caller has drained the stream).
void drelease(FusedBufs &B, void *p) {
    auto it = std::find(B.own.begin(), B.own.end(), p);
    if (it == B.own.end()) return;
    (void)hipFree(p);
    B.own.erase(it);
}

}  // namespace

int fused_setup(tdt_ctx *c) {
    if (c->fused) return TDT_OK;
    FusedBufs *B = new FusedBufs();
    const Dev &D = c->D;
    FDev &F = B->F;
    F.n = D.n; F.g = D.g; F.m = D.m; F.nt = D.nt;
    F.remap = 1;
    F.nc4 = D.nc4; F.nod4 = D.nod4; F.xn = D.xn; F.yn = D.yn;
    F.gparent = D.gparent; F.gtype = D.gtype; F.gopp = D.gopp; F.gside = D.gside;
    F.g6ptr = D.g6ptr; F.g6cell = D.g6cell; F.g6dely = D.g6dely;
    F.has_bc6 = !c->pm.g6cell.empty();
    F.nptr = D.nptr; F.nlist = D.nlist; F.novr = D.novr;
    F.xcc = D.xcc; F.ycc = D.ycc; F.vol = D.vol; F.dl = D.dl; F.sc = D.sc;
    F.gxcc = D.xcc; F.gycc = D.ycc; F.gvol = D.vol; F.gsc = D.sc; F.gxn = D.xn; F.gyn = D.yn;
    F.qn = D.c;
    F.dt = D.dt;
    bool ok = true;
    if (c->precision == TDT_FP64) {
        // the routine-level arrays double as the fused path's storage
        // the routine-level D and VIS arrays double as the fused path's
        F.dd = D.d24; F.vis = D.vis;
        F.drec = dalloc<double>(*B, (size_t)kRecLen * D.n);
        F.nrec = dalloc<double>(*B, (size_t)5 * D.m);
        F.nxy = dalloc<double>(*B, (size_t)2 * D.m);
        F.st[0] = dalloc<double>(*B, (size_t)5 * D.n);
        F.st[1] = dalloc<double>(*B, (size_t)5 * D.n);
        ok = F.drec && F.nrec && F.nxy && F.st[0] && F.st[1];
    } else {
        F.drec = dalloc<float>(*B, (size_t)kRecLen * D.n);
        F.nrec = dalloc<float>(*B, (size_t)5 * D.m);
        F.nxy = dalloc<float>(*B, (size_t)2 * D.m);
        F.dd = dalloc<float>(*B, (size_t)5 * D.n);
        F.vis = dalloc<float>(*B, (size_t)4 * D.n);
        F.st[0] = dalloc<float>(*B, (size_t)5 * D.n);
        F.st[1] = dalloc<float>(*B, (size_t)5 * D.n);
        ok = F.drec && F.nrec && F.nxy && F.dd && F.vis && F.st[0] && F.st[1];
        F.gxcc = dalloc<float>(*B, D.nt);
        F.gycc = dalloc<float>(*B, D.nt);
        F.gvol = dalloc<float>(*B, D.nt);
        F.gsc = dalloc<float>(*B, (size_t)8 * D.n);
        F.gxn = dalloc<float>(*B, D.m);
        F.gyn = dalloc<float>(*B, D.m);
        ok = ok && F.gxcc && F.gycc && F.gvol && F.gsc && F.gxn && F.gyn;
    }
    if (F.has_bc6) {
        F.g6s = dalloc<double>(*B, (size_t)6 * D.g);
        ok = ok && F.g6s;
    }
    B->nb_cells = (int)nblocks(D.n);
    F.ncb8 = (B->nb_cells + 7) & ~7;
    B->nb_nodes = (int)nblocks(D.m);
    if (c->batch) {
        // block tables: every block holds cells (nodes) of one case
        const BatchInfo &bi = *c->batch;
        F.ncases = bi.ncases;
        F.cblk = bi.d_cblk;
        F.nblk = bi.d_nblk;
        F.ccase = bi.d_ccase;
        F.kc64 = bi.d_k64;
        F.kcr = c->precision == TDT_FP64 ? (const void *)bi.d_k64 : (const void *)bi.d_kf;
        B->nb_cells = bi.nb_cells;
        F.ncb8 = (int)bi.cblk.size();
        B->nb_nodes = bi.nb_nodes;
    }
    B->nbq_cells = (D.n + kBlock / 4 - 1) / (kBlock / 4);
    B->nbq_nodes = (D.m + kBlock / 4 - 1) / (kBlock / 4);
    // TDT_KERNELS=merged|split|quad (or a kernel-set number) pins the kernel
    // set; otherwise the first tdt_iterate call picks it by timing (autotune)
    const char *kv = std::getenv("TDT_KERNELS");
    const std::string ks = kv ? kv : "";
    // the face-once viscous set exists for FP32-mixed single cases (in FP64 every
    // set must give the reference's per-cell evaluation, bit for bit)
    const bool face_ok = c->precision != TDT_FP64 && !c->batch;
    if (ks == "merged" || ks == "split" || ks == "quad" || (ks == "face" && face_ok)) {
        B->kset = ks == "merged" ? kSetMerged : ks == "split" ? kSetSplit
                  : ks == "quad" ? kSetQuad : kSetFace;
        B->tuned = true;
    } else if (!ks.empty() && ks.find_first_not_of("0123456789") == std::string::npos) {
        const int v = std::atoi(ks.c_str());
        if ((v & ~kSetBits) == 0 && ((v & kSt1Mask) != kSt1Face || face_ok)) {
            B->kset = v;
            B->tuned = true;
        }
    }
    if (c->batch) B->kset &= kSt1Split | kStPreload;   // no quad form for a batch
    // the wave node tables: FP64 single cases with LOCAL = 1 only
    if (c->precision != TDT_FP64 || c->batch || !c->K.local) B->kset &= ~kDvNodeLds;
    // fused-only neighbour and node-gather tables (built from the prepared mesh)
    const PreparedMesh &pm = c->pm;
    if (D.nt > kIdMask) {
        c->fused = B;
        fused_free(c);
        c->err = "fused path: more than 2^28 cells and ghosts";
        return TDT_ERR_UNSUPPORTED;
    }
    std::vector<int> nct(pm.nc4);
    for (int &v : nct)
        if (v >= D.n) v |= pm.gtype[v - D.n] << kTypeShift;
    std::vector<int> ncon((size_t)2 * kNodeWidth * D.m, 0), nrel(D.m, 0);
    for (int v = 0; v < D.m; ++v) {
        const int j0 = pm.nptr[v], cnt = pm.nptr[v + 1] - j0;
        int *e = &ncon[(size_t)2 * kNodeWidth * v];
        for (int k = 0; k < kNodeWidth && k < cnt; ++k) {
            const int id = pm.nlist[j0 + k];
            int a = id, src = id;
            if (id >= D.n) {
                const int g = id - D.n, ty = pm.gtype[g];
                src = pm.gparent[g];
                a = (ty == 3 ? pm.gopp[g] : id) | ty << kTypeShift;
            }
            e[2 * k] = a;
            e[2 * k + 1] = src;
        }
        for (int k = cnt; k < kNodeWidth && cnt > 0; ++k) {   // padding, unused
            e[2 * k] = e[0];
            e[2 * k + 1] = e[1];
        }
        // sum starts relative to the node's first contributor (0..cnt)
        const int su = pm.novr[4 * v] - j0, sv = pm.novr[4 * v + 1] - j0,
                  sr = pm.novr[4 * v + 2] - j0;
        nrel[v] = cnt <= kNodeWidth
                      ? cnt | su << 4 | sv << 8 | sr << 12 | (pm.novr[4 * v + 3] ? 1 << 16 : 0)
                      : 15;   // more contributors than the records hold: list path
    }
    if (face_ok) {
        // viscous faces: each face once, owned by the lower cell id (a boundary
        // face by its cell), in owner order; the non-owner finds it by its nodes
        std::vector<int4> vfl;
        std::vector<int> cvf((size_t)4 * D.n, -1);
        vfl.reserve((size_t)2 * D.n + D.g);
        for (int i = 0; i < D.n; ++i)
            for (int f = 0; f < 4; ++f) {
                const int j = pm.nc4[4 * i + f];
                if (j < D.n && j <= i) continue;
                cvf[4 * i + f] = (int)vfl.size();
                vfl.push_back(make_int4(i | f << kTypeShift, nct[4 * i + f], pm.nod4[4 * i + f],
                                        pm.nod4[4 * i + ((f + 1) & 3)]));
            }
        for (int i = 0; i < D.n && ok; ++i)
            for (int f = 0; f < 4; ++f) {
                if (cvf[4 * i + f] >= 0) continue;
                const int j = pm.nc4[4 * i + f];
                const int b = pm.nod4[4 * i + f], d = pm.nod4[4 * i + ((f + 1) & 3)];
                int hit = -1;
                for (int fj = 0; fj < 4 && hit < 0; ++fj) {
                    const int e = cvf[4 * j + fj];
                    if (e >= 0 && pm.nc4[4 * j + fj] == i && vfl[e].z == d && vfl[e].w == b)
                        hit = e;
                }
                if (hit < 0) {
                    ok = false;   // no matching face: the grid is not a conforming quad mesh
                    break;
                }
                cvf[4 * i + f] = hit | (int)0x80000000u;
            }
        F.nf = (int)vfl.size();
        F.nfb8 = ((int)nblocks(F.nf) + 7) & ~7;
        int4 *d_vfl = dalloc<int4>(*B, vfl.size());
        int4 *d_cvf = dalloc<int4>(*B, D.n);
        F.vflux = dalloc<float4>(*B, vfl.size());
        ok = ok && d_vfl && d_cvf && F.vflux &&
             hipMemcpy(d_vfl, vfl.data(), vfl.size() * sizeof(int4), hipMemcpyHostToDevice) ==
                 hipSuccess &&
             hipMemcpy(d_cvf, cvf.data(), cvf.size() * sizeof(int), hipMemcpyHostToDevice) ==
                 hipSuccess;
        F.vfl = d_vfl;
        F.cvf = d_cvf;
    }
    if (c->precision == TDT_FP64 && !c->batch && ok) {
        // node tables of the wavefronts (kDvNodeLds)
        std::vector<int> bnode, nloc;
        const int cap = wave_node_tables(pm.nod4, D.n, B->nb_cells * (kBlock / kWave), bnode,
                                         nloc);
        int *d_bnode = dalloc<int>(*B, bnode.size());
        int *d_nloc = dalloc<int>(*B, nloc.size());
        ok = d_bnode && d_nloc &&
             hipMemcpy(d_bnode, bnode.data(), bnode.size() * sizeof(int),
                       hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(d_nloc, nloc.data(), nloc.size() * sizeof(int), hipMemcpyHostToDevice) ==
                 hipSuccess;
        F.bnode = d_bnode;
        F.nloc = reinterpret_cast<const int2 *>(d_nloc);
        F.nl_cap = cap;
    }
    int *d_nct = dalloc<int>(*B, nct.size());
    int *d_ncon = dalloc<int>(*B, ncon.size());
    int *d_nrel = dalloc<int>(*B, nrel.size());
    ok = ok && d_nct && d_ncon && d_nrel &&
         hipMemcpy(d_nct, nct.data(), nct.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(d_ncon, ncon.data(), ncon.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(d_nrel, nrel.data(), nrel.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    F.nct = reinterpret_cast<const int4 *>(d_nct);
    F.ncon = reinterpret_cast<const int2 *>(d_ncon);
    F.nrel = d_nrel;
    B->qbak = dalloc<double>(*B, (size_t)5 * D.nt);
    F.dtslot = dalloc<unsigned long long>(*B, 1);
    F.ctr = dalloc<int>(*B, 4);
    F.part = dalloc<double>(*B, (size_t)std::max(B->nb_cells, B->nbq_cells) * kPartStride);
    ok = ok && F.dtslot && F.ctr && F.part && B->qbak;
    if (ok)
        for (int k = 0; k < 12; ++k) ok = ok && hipEventCreate(&B->ev[k]) == hipSuccess;
    B->have_events = ok;
    if (!ok) {
        c->fused = B;
        fused_free(c);
        c->err = "device allocation for the fused path failed";
        return TDT_ERR_DEVICE;
    }
    c->fused = B;
    return TDT_OK;
}

void fused_free(tdt_ctx *c) {
    FusedBufs *B = c->fused;
    if (!B) return;
    if (B->exec1) hipGraphExecDestroy(B->exec1);
    if (B->execK) hipGraphExecDestroy(B->execK);
    if (B->graph1) hipGraphDestroy(B->graph1);
    if (B->graphK) hipGraphDestroy(B->graphK);
    if (B->have_events)
        for (int k = 0; k < 12; ++k) hipEventDestroy(B->ev[k]);
    for (void *p : B->own) hipFree(p);
    if (B->hist_host) hipHostFree(B->hist_host);
    delete B;
    c->fused = nullptr;
}

static hipError_t build_graph(tdt_ctx *c, FusedBufs &B, int iters, hipGraph_t *g,
                              hipGraphExec_t *ge) {
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    int prof = c->profiling;
    c->profiling = 0;
    for (int k = 0; k < iters && e == hipSuccess; ++k) e = enqueue(c, B);
    c->profiling = prof;
    hipError_t e2 = hipStreamEndCapture(c->stream, g);
    if (e != hipSuccess) return e;
    if (e2 != hipSuccess) return e2;
    return hipGraphInstantiate(ge, *g, nullptr, nullptr, 0);
}

static void drop_graphs(FusedBufs &B) {
    if (B.exec1) { (void)hipGraphExecDestroy(B.exec1); B.exec1 = nullptr; }
    if (B.execK) { (void)hipGraphExecDestroy(B.execK); B.execK = nullptr; }
    if (B.graph1) { (void)hipGraphDestroy(B.graph1); B.graph1 = nullptr; }
    if (B.graphK) { (void)hipGraphDestroy(B.graphK); B.graphK = nullptr; }
}

// The graphs hold the flow constants as kernel arguments: after a change of
// RINIT (RTOL) they are captured again on the next iterate.
void fused_drop_graphs(tdt_ctx *c) {
    if (!c->fused) return;
    (void)hipStreamSynchronize(c->stream);
    drop_graphs(*c->fused);
}

// Runs `iters` iterations of kernel set `kset` eagerly.  With per_slot, each
// kernel is bracketed by events and per_slot[q] receives the average time of
// slot q; otherwise *total_ms receives the whole run's time.
static hipError_t time_set(tdt_ctx *c, FusedBufs &B, int kset, int iters, double *per_slot,
                           float *total_ms) {
    const FDev &F = B.F;
    B.kset = kset;
    c->profiling = per_slot ? 1 : 0;
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, c->stream, F.ctr, F.dtslot);
    hipError_t e = enqueue(c, B);                                   // warm-up
    if (e == hipSuccess && per_slot) {
        e = hipEventSynchronize(B.ev[2 * kNumFusedKernels - 1]);
        for (int q = 0; q < kNumFusedKernels; ++q) per_slot[q] = 0.0;
        for (int it = 0; it < iters && e == hipSuccess; ++it) {
            e = enqueue(c, B);
            if (e == hipSuccess) e = hipEventSynchronize(B.ev[2 * kNumFusedKernels - 1]);
            for (int q = 0; q < kNumFusedKernels && e == hipSuccess; ++q) {
                float ms = 0.f;
                e = hipEventElapsedTime(&ms, B.ev[2 * q], B.ev[2 * q + 1]);
                per_slot[q] += ms / iters;
            }
        }
    } else if (e == hipSuccess) {
        e = hipEventRecord(B.ev[10], c->stream);
        for (int r = 0; r < iters && e == hipSuccess; ++r) e = enqueue(c, B);
        if (e == hipSuccess) e = hipEventRecord(B.ev[11], c->stream);
        if (e == hipSuccess) e = hipEventSynchronize(B.ev[11]);
        if (e == hipSuccess) e = hipEventElapsedTime(total_ms, B.ev[10], B.ev[11]);
    }
    c->profiling = 0;
    return e;
}

// Picks the kernel set by timing on the current state.  The uniform sets
// (merged, split, quad) are timed kernel by kernel; the fastest choice for each
// slot (pre-pass, stage 1, stages 2-4) makes a candidate, which then competes
// with the uniform sets on whole iterations.  Everything an iteration reads
// besides Q^n is recomputed within it, so restoring Q^n after each trial leaves
// the solve exactly where it was; in FP64 every set computes the same bits.
static hipError_t autotune(tdt_ctx *c, FusedBufs &B) {
    const FDev &F = B.F;
    const size_t qbytes = (size_t)5 * F.nt * sizeof(double);
    const int prof = c->profiling;
    hipError_t e = hipMemcpyAsync(B.qbak, F.qn, qbytes, hipMemcpyDeviceToDevice, c->stream);
    const int sets[4] = {kSetMerged, kSetSplit, kSetQuad, kSetFace};
    // a batch has no quad form; the face set is FP32-mixed single cases only
    const int nsets = F.ncases ? 2 : F.nf > 0 ? 4 : 3;
    double t[4][kNumFusedKernels] = {};
    for (int k = 0; k < nsets && e == hipSuccess; ++k) {
        e = time_set(c, B, sets[k], kTuneIterations, t[k], nullptr);
        if (e == hipSuccess)
            e = hipMemcpyAsync(F.qn, B.qbak, qbytes, hipMemcpyDeviceToDevice, c->stream);
    }
    // per slot: pre-pass (merged's k_pre vs quad's kq_pre), stage 1 (merged,
    // split = k_dv + stage 1, quad), stages 2-4 (merged's vs quad's)
    int comp = nsets > 2 && t[2][0] < t[0][0] ? kPreQuad : 0;
    const double s1[4] = {t[0][2], t[1][1] + t[1][2], t[2][2], t[3][1] + t[3][2]};
    const int s1set[4] = {0, kSt1Split, kSt1Quad, kSt1Face};
    int b1 = 0;
    for (int k = 1; k < nsets; ++k) b1 = s1[k] < s1[b1] ? k : b1;
    comp |= s1set[b1];
    if (nsets > 2 && t[2][3] + t[2][4] + t[2][5] < t[0][3] + t[0][4] + t[0][5]) comp |= kStQuad;
    // whole iterations: the composed set against the uniform ones
    const int cand[5] = {kSetMerged, kSetSplit, kSetQuad, kSetFace, comp};
    int best = kSetMerged;
    float best_ms = 3.4e38f;
    for (int k = 0; k < 5 && e == hipSuccess; ++k) {
        if ((k == 2 && nsets < 3) || (k == 3 && nsets < 4)) continue;
        if (k == 4 && (comp == kSetMerged || comp == kSetSplit || comp == kSetQuad ||
                       comp == kSetFace))
            break;
        float ms = 0.f;
        e = time_set(c, B, cand[k], kTuneIterations, nullptr, &ms);
        if (e == hipSuccess)
            e = hipMemcpyAsync(F.qn, B.qbak, qbytes, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess && ms < best_ms) { best_ms = ms; best = cand[k]; }
    }
    // FP64: the one-lane stage kernels with neighbour states preloaded, if the
    // chosen set has any; then k_dv with the node tables in LDS, if it has that
    if (e == hipSuccess && c->precision == TDT_FP64 &&
        ((best & kSt1Mask) == kSt1Split || !(best & kStQuad))) {
        float ms = 0.f;
        e = time_set(c, B, best | kStPreload, kTuneIterations, nullptr, &ms);
        if (e == hipSuccess)
            e = hipMemcpyAsync(F.qn, B.qbak, qbytes, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess && ms < best_ms) {
            best_ms = ms;
            best = best | kStPreload;
        }
    }
    if (e == hipSuccess && (best & kSt1Mask) == kSt1Split && node_lds_usable(c, B)) {
        float ms = 0.f;
        e = time_set(c, B, best | kDvNodeLds, kTuneIterations, nullptr, &ms);
        if (e == hipSuccess)
            e = hipMemcpyAsync(F.qn, B.qbak, qbytes, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess && ms < best_ms) best = best | kDvNodeLds;
    }
    c->profiling = prof;
    B.kset = best;
    B.tuned = true;
    drop_graphs(B);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
}

int fused_kernel_set(tdt_ctx *c) { return c->fused ? c->fused->kset : -1; }

int fused_iterate(tdt_ctx *c, int niter, double *l2, double *linf, int *first_nonfinite) {
    int rc = fused_setup(c);
    if (rc != TDT_OK) return rc;
    FusedBufs &B = *c->fused;
    FDev &F = B.F;
    if (B.geom_version != c->geom_version) {
        // FP32 copies of the geometry the flux arithmetic reads, and the node
        // coordinate pairs the viscous faces read
        const Dev &D = c->D;
        const bool fp64 = c->precision == TDT_FP64;
        const struct { const double *s; const void *d; int cnt; } cp[] = {
            {D.xcc, F.gxcc, D.nt}, {D.ycc, F.gycc, D.nt}, {D.vol, F.gvol, D.nt},
            {D.sc, F.gsc, 8 * D.n}, {D.xn, F.gxn, D.m}, {D.yn, F.gyn, D.m}};
        for (const auto &x : cp)
            if (!fp64)
                hipLaunchKernelGGL(k_narrow, dim3(nblocks(x.cnt)), dim3(kBlock), 0, c->stream,
                                   x.s, (float *)x.d, x.cnt);
        if (fp64)
            hipLaunchKernelGGL(k_node_xy<double>, dim3(nblocks(D.m)), dim3(kBlock), 0, c->stream,
                               F.gxn, F.gyn, F.nxy, D.m);
        else
            hipLaunchKernelGGL(k_node_xy<float>, dim3(nblocks(D.m)), dim3(kBlock), 0, c->stream,
                               F.gxn, F.gyn, F.nxy, D.m);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            c->err = std::string("geometry copies: ") + hipGetErrorString(e);
            return TDT_ERR_DEVICE;
        }
        // FP64 kernels that gather a face's nodes form its surface vector from
        // them while SC is the nodes' (the graphs hold F: captured again)
        const int scn = fp64 && c->sc_from_nodes ? (c->params.iaxis == 1 ? 2 : 1) : 0;
        if (scn != F.sc_nodes) {
            (void)hipStreamSynchronize(c->stream);
            drop_graphs(B);
            F.sc_nodes = scn;
        }
        B.geom_version = c->geom_version;
    }
    const size_t nrec = F.ncases ? (size_t)F.ncases : 1;   // history records per iteration
    if ((size_t)niter > B.hist_cap) {
        // room for 65536 iterations of one case (5 MB) from the first call, so
        // runs of usual lengths never reallocate it and recapture their graphs
        size_t cap = std::max((size_t)niter, std::max((size_t)1024, (size_t)65536 / nrec));
        double *h = dalloc<double>(B, cap * 10 * nrec);
        if (!h) {
            c->err = "device allocation of the monitor history failed";
            return TDT_ERR_DEVICE;
        }
        // the history is read back into pinned memory: a pageable copy of this
        // size pays a one-time runtime staging cost of several milliseconds
        double *hh = nullptr;
        if (hipHostMalloc((void **)&hh, cap * 10 * nrec * sizeof(double), 0) != hipSuccess) {
            c->err = "pinned allocation of the monitor history failed";
            return TDT_ERR_DEVICE;
        }
        // the old buffers: the stream has no work left that reads them
        hipError_t es = hipStreamSynchronize(c->stream);
        if (es != hipSuccess) {
            c->err = std::string("monitor history: ") + hipGetErrorString(es);
            return TDT_ERR_DEVICE;
        }
        if (B.hist_host) (void)hipHostFree(B.hist_host);
        if (F.hist) drelease(B, F.hist);
        B.hist_host = hh;
        F.hist = h;
        B.hist_cap = cap;
        // one copy of the whole buffer now: the runtime sets up its path for
        // large device-to-host copies on first use (several milliseconds), which
        // would otherwise land in the first run that reads a long history back
        hipError_t ec = hipMemcpyAsync(hh, h, cap * 10 * nrec * sizeof(double),
                                       hipMemcpyDeviceToHost, c->stream);
        if (ec == hipSuccess) ec = hipStreamSynchronize(c->stream);
        if (ec != hipSuccess) {
            c->err = std::string("monitor history: ") + hipGetErrorString(ec);
            return TDT_ERR_DEVICE;
        }
        // kernels captured earlier point at the old history buffer
        if (B.exec1) { hipGraphExecDestroy(B.exec1); B.exec1 = nullptr; }
        if (B.execK) { hipGraphExecDestroy(B.execK); B.execK = nullptr; }
        if (B.graph1) { hipGraphDestroy(B.graph1); B.graph1 = nullptr; }
        if (B.graphK) { hipGraphDestroy(B.graphK); B.graphK = nullptr; }
    }
    hipError_t e = hipSuccess;
    if (!B.tuned && niter > 0) {
        e = autotune(c, B);
        if (e != hipSuccess) {
            c->err = std::string("autotune: ") + hipGetErrorString(e);
            return TDT_ERR_DEVICE;
        }
    }
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, c->stream, F.ctr, F.dtslot);
    e = hipGetLastError();
    for (int q = 0; q < kNumFusedKernels; ++q) c->ktimes[q] = 0.0;
    if (e == hipSuccess && c->profiling) {
        // eager launches bracketed by events: per-kernel device time
        for (int it = 0; it < niter && e == hipSuccess; ++it) {
            e = enqueue(c, B);
            if (e == hipSuccess) e = hipEventSynchronize(B.ev[2 * kNumFusedKernels - 1]);
            for (int q = 0; q < kNumFusedKernels && e == hipSuccess; ++q) {
                float ms = 0.f;
                e = hipEventElapsedTime(&ms, B.ev[2 * q], B.ev[2 * q + 1]);
                c->ktimes[q] += ms;
            }
        }
        for (int q = 0; q < kNumFusedKernels; ++q) c->ktimes[q] /= (niter > 0 ? niter : 1);
        if ((B.kset & kSt1Mask) != kSt1Split && (B.kset & kSt1Mask) != kSt1Face)
            c->ktimes[1] = 0.0;   // no k_dv / k_vface launch
    } else if (e == hipSuccess) {
        // both graphs are captured on the first call, whatever its length, so
        // that later calls only replay them
        if (!B.exec1) e = build_graph(c, B, 1, &B.graph1, &B.exec1);
        if (e == hipSuccess && !B.execK) e = build_graph(c, B, B.chunk, &B.graphK, &B.execK);
        int done = 0;
        while (e == hipSuccess && niter - done >= B.chunk && B.execK) {
            e = hipGraphLaunch(B.execK, c->stream);
            done += B.chunk;
        }
        while (e == hipSuccess && done < niter) {
            e = hipGraphLaunch(B.exec1, c->stream);
            ++done;
        }
    }
    if (e == hipSuccess && F.ncases) {
        hipLaunchKernelGGL(k_finalize_batch, dim3(F.ncases), dim3(kBlock), 0, c->stream, F);
        hipLaunchKernelGGL(k_batch_tick, dim3(1), dim3(1), 0, c->stream, F.ctr);
    } else if (e == hipSuccess) {
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, c->stream, F,
                           partial_blocks(B, B.kset));
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_cold_from_qn, dim3(nblocks(F.n)), dim3(kBlock), 0, c->stream, F,
                           c->D.cold);
    }
    if (e == hipSuccess) e = hipGetLastError();
    int ctr[4] = {0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(ctr, F.ctr, sizeof ctr, hipMemcpyDeviceToHost, c->stream);
    const size_t nvals = (l2 || linf) && niter > 0 ? (size_t)niter * 10 * nrec : 0;
    if (e == hipSuccess && nvals)
        e = hipMemcpyAsync(B.hist_host, F.hist, nvals * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        c->err = std::string("fused iteration: ") + hipGetErrorString(e);
        return TDT_ERR_DEVICE;
    }
    // records [iteration][case]: L2 in 0..4, Linf in 5..9
    const double *h = B.hist_host;
    for (size_t r = 0; r < nvals / 10; ++r)
        for (int q = 0; q < 5; ++q) {
            if (l2) l2[r * 5 + q] = h[r * 10 + q];
            if (linf) linf[r * 5 + q] = h[r * 10 + 5 + q];
        }
    if (first_nonfinite) *first_nonfinite = ctr[2];
    if (ctr[0] != niter) {
        c->err = "fused iteration: device iteration counter mismatch";
        return TDT_ERR_DEVICE;
    }
    if (ctr[2]) {
        c->err = "non-finite value in the state at iteration " + std::to_string(ctr[2]);
        return TDT_ERR_NONFINITE;
    }
    return TDT_OK;
}

}  // namespace tdt
