// C ABI of libtdt, batches of cases (tdt_batch_*, include/tdt.h): independent
// grids concatenated into one context and advanced by one fused iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "api_internal.hpp"

using tdt::Consts;
using namespace tdt::api;

namespace {

// Batch: the cases' flow constants with their RTOL (set_rinit's arithmetic),
// uploaded for the kernels.
int batch_upload_consts(tdt_ctx *c) {
    tdt::BatchInfo &bi = *c->batch;
    for (int k = 0; k < bi.ncases; ++k) {
        const double rtred = bi.params[k].rtred;
        const double rtol = bi.rinit[k] * rtred;
        bi.k64[k].rtol = rtol;
        bi.k64[k].rtol_rtred = rtol * rtred;
        bi.kf[k].rtol = (float)rtol;
        bi.kf[k].rtol_rtred = (float)bi.k64[k].rtol_rtred;
    }
    HIPCALL(c, hipMemcpyAsync(bi.d_k64, bi.k64.data(), bi.k64.size() * sizeof(Consts<double>),
                              hipMemcpyHostToDevice, c->stream), "batch constants");
    HIPCALL(c, hipMemcpyAsync(bi.d_kf, bi.kf.data(), bi.kf.size() * sizeof(Consts<float>),
                              hipMemcpyHostToDevice, c->stream), "batch constants");
    return sync(c);
}

// Batch: case icase's cells (internal range) to or from host order.
int batch_case_range(tdt_ctx *c, int icase, int *first, int *count) {
    if (!c->batch) {
        c->err = "not a batch context";
        return TDT_ERR_ARG;
    }
    if (icase < 0 || icase >= c->batch->ncases) {
        c->err = "case index out of range";
        return TDT_ERR_ARG;
    }
    *first = c->batch->cell_off[icase];
    *count = c->batch->ncell[icase];
    return TDT_OK;
}

}  // namespace

extern "C" {

// ---- batched cases ----------------------------------------------------------------

int tdt_batch_create(tdt_ctx **out, int32_t ncases, const tdt_mesh *meshes,
                     const tdt_params *params, int32_t precision, int32_t ordering) {
    g_create_error.clear();
    if (!out || !meshes || !params || ncases < 1) {
        g_create_error = "null argument or no cases";
        return TDT_ERR_ARG;
    }
    *out = nullptr;
    if (precision != TDT_FP64 && precision != TDT_FP32_MIXED) {
        g_create_error = "unknown precision";
        return TDT_ERR_ARG;
    }
    for (int k = 0; k < ncases; ++k) {
        if (!params_ok(params[k])) {
            g_create_error = "case " + std::to_string(k) + ": " + g_create_error;
            return TDT_ERR_ARG;
        }
        // per case: P0, T0, AMACHINIT, PAMB, RTRED and the BC type-7 inputs; the
        // gas and model constants are the batch's
        const tdt_params &a = params[k], &b = params[0];
        if (a.local != 1 || a.iaxis != b.iaxis || a.gama != b.gama || a.amol != b.amol ||
            a.cfl != b.cfl || a.prandtl != b.prandtl || a.prandtlt != b.prandtlt ||
            a.npseudo != b.npseudo) {
            g_create_error = "case " + std::to_string(k) +
                             ": a batch needs LOCAL = 1 and the same IAXIS, GAMA, AMOL, CFL, "
                             "PRANDTL, PRANDTLT and NPSEUDO in every case";
            return TDT_ERR_UNSUPPORTED;
        }
    }
    // each case prepared alone: its validation and the cell order it would get alone
    auto bi = std::make_unique<tdt::BatchInfo>();
    bi->ncases = ncases;
    std::vector<int> order;
    size_t N = 0, G = 0, M = 0;
    for (int k = 0; k < ncases; ++k) {
        const tdt_mesh &ms = meshes[k];
        tdt::PreparedMesh pk;
        int code = TDT_OK;
        std::string msg = tdt::prepare_mesh(ms, ordering, pk, &code);
        if (!msg.empty()) {
            g_create_error = "case " + std::to_string(k) + ": " + msg;
            return code;
        }
        if (!pk.g6cell.empty()) {
            g_create_error = "case " + std::to_string(k) + ": BC type 6 is not available in a batch";
            return TDT_ERR_UNSUPPORTED;
        }
        if (N + ms.neles + G + ms.nghosts > ((size_t)1 << 28)) {   // the fused path's id range
            g_create_error = "batch: too many cells and ghosts";
            return TDT_ERR_UNSUPPORTED;
        }
        bi->cell_off.push_back((int)N);
        bi->ncell.push_back(ms.neles);
        bi->ghost_off.push_back((int)G);
        bi->nghost.push_back(ms.nghosts);
        bi->orig_cell_off.push_back((int)N);
        for (int q : pk.cell_new2old) order.push_back((int)N + q);
        N += ms.neles;
        G += ms.nghosts;
        M += ms.nods;
    }
    // the concatenated grid: cells, then ghosts, of case 0, 1, ... (grid.in ids offset)
    std::vector<double> x(M), y(M);
    std::vector<int32_t> nod(4 * N), nc(4 * N), par(G), gh(G), ty(G), sd(G);
    size_t moff = 0;
    for (int k = 0; k < ncases; ++k) {
        const tdt_mesh &ms = meshes[k];
        const int n = ms.neles, noff = bi->cell_off[k], goff = bi->ghost_off[k];
        std::unordered_map<int, int> gpos;
        for (int j = 0; j < ms.nghosts; ++j) gpos[ms.ghost[j]] = j;
        for (int v = 0; v < ms.nods; ++v) {
            x[moff + v] = ms.x_mm[v];
            y[moff + v] = ms.y_mm[v];
        }
        for (int f = 0; f < 4; ++f)
            for (int i = 0; i < n; ++i) {
                const size_t s = (size_t)f * n + i, d = (size_t)f * N + noff + i;
                nod[d] = (int32_t)(moff + ms.nod[s]);
                const int v = ms.nc[s];
                nc[d] = v >= 1 && v <= n ? noff + v
                        : v == -1        ? -1
                                         : (int32_t)(N + goff + gpos.at(v) + 1);
            }
        for (int j = 0; j < ms.nghosts; ++j) {
            par[goff + j] = noff + ms.parent[j];
            gh[goff + j] = (int32_t)(N + goff + j + 1);
            ty[goff + j] = ms.type[j];
            sd[goff + j] = ms.side[j];
        }
        moff += ms.nods;
    }
    tdt_mesh all{};
    all.nods = (int32_t)M;
    all.neles = (int32_t)N;
    all.nghosts = (int32_t)G;
    all.x_mm = x.data();
    all.y_mm = y.data();
    all.nod = nod.data();
    all.nc = nc.data();
    all.parent = par.data();
    all.ghost = gh.data();
    all.type = ty.data();
    all.side = sd.data();
    tdt_ctx *c = new tdt_ctx();
    c->precision = precision;
    int code = TDT_OK;
    std::string msg = tdt::prepare_mesh(all, TDT_ORDER_INPUT, c->pm, &code, &order);
    if (!msg.empty()) {
        g_create_error = "batch: " + msg;
        delete c;
        return code;
    }
    code = init_device(c, &params[0]);
    if (code != TDT_OK) {
        tdt_destroy(c);
        return code;
    }
    set_rinit(c, 0.0);
    // cell blocks: each holds cells of one case
    for (int k = 0; k < ncases; ++k) {
        const int first = (int)bi->cblk.size(), n = bi->ncell[k];
        for (int b = 0; b * tdt::kBlock < n; ++b)
            bi->cblk.push_back(make_int4(bi->cell_off[k] + b * tdt::kBlock,
                                         std::min(tdt::kBlock, n - b * tdt::kBlock), k, 0));
        bi->ccase.push_back(make_int4(first, (int)bi->cblk.size() - first, n, 0));
    }
    bi->nb_cells = (int)bi->cblk.size();
    while (bi->cblk.size() % 8) bi->cblk.push_back(make_int4(0, 0, 0, 0));
    // node blocks: a node belongs to the case of the first cell that touches it
    // (nodes no cell touches come last; they are given case 0)
    const tdt::PreparedMesh &pm = c->pm;
    std::vector<int> ncase(pm.m, -1);
    for (int k = 0; k < ncases; ++k)
        for (int i = bi->cell_off[k]; i < bi->cell_off[k] + bi->ncell[k]; ++i)
            for (int f = 0; f < 4; ++f) {
                int &v = ncase[pm.nod4[(size_t)4 * i + f]];
                if (v < 0) v = k;
            }
    for (int v = 0; v < pm.m;) {
        const int k = ncase[v] < 0 ? 0 : ncase[v];
        int e = v + 1;
        while (e < pm.m && e - v < tdt::kBlock && (ncase[e] < 0 ? 0 : ncase[e]) == k) ++e;
        bi->nblk.push_back(make_int4(v, e - v, k, 0));
        v = e;
    }
    bi->nb_nodes = (int)bi->nblk.size();
    while (bi->nblk.size() % 8) bi->nblk.push_back(make_int4(0, 0, 0, 0));
    for (int k = 0; k < ncases; ++k) {
        bi->params.push_back(params[k]);
        bi->k64.push_back(make_consts<double>(params[k]));
        bi->kf.push_back(make_consts<float>(params[k]));
    }
    bi->rinit.assign(ncases, 0.0);
    bi->d_k64 = dev_alloc<Consts<double>>(c, ncases);
    bi->d_kf = dev_alloc<Consts<float>>(c, ncases);
    bi->d_cblk = dev_alloc<int4>(c, bi->cblk.size());
    bi->d_nblk = dev_alloc<int4>(c, bi->nblk.size());
    bi->d_ccase = dev_alloc<int4>(c, ncases);
    bool ok = bi->d_k64 && bi->d_kf && bi->d_cblk && bi->d_nblk && bi->d_ccase;
    auto up = [&](void *dst, const void *src, size_t bytes) {
        if (ok) ok = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    up(bi->d_cblk, bi->cblk.data(), bi->cblk.size() * sizeof(int4));
    up(bi->d_nblk, bi->nblk.data(), bi->nblk.size() * sizeof(int4));
    up(bi->d_ccase, bi->ccase.data(), bi->ccase.size() * sizeof(int4));
    c->batch = bi.release();
    if (!ok || batch_upload_consts(c) != TDT_OK) {
        g_create_error = "batch: device allocation or upload failed";
        tdt_destroy(c);
        return TDT_ERR_DEVICE;
    }
    *out = c;
    return TDT_OK;
}

int tdt_batch_cold_start(tdt_ctx *c, double *rinit) {
    CHECK_CTX(c);
    if (!c->batch) {
        c->err = "not a batch context";
        return TDT_ERR_ARG;
    }
    tdt::BatchInfo &bi = *c->batch;
    const size_t n = c->pm.n, nt = c->D.nt;
    std::vector<double> h(5 * n);
    for (int k = 0; k < bi.ncases; ++k) {
        double q[5];
        cold_state(bi.params[k], bi.k64[k], q, &bi.rinit[k]);
        for (int m = 0; m < 5; ++m)
            std::fill(h.begin() + m * n + bi.cell_off[k],
                      h.begin() + m * n + bi.cell_off[k] + bi.ncell[k], q[m]);
        if (rinit) rinit[k] = bi.rinit[k];
    }
    for (int m = 0; m < 5; ++m) {
        HIPCALL(c, hipMemcpyAsync(c->D.c + m * nt, h.data() + m * n, n * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream), "batch cold start");
        HIPCALL(c, hipMemcpyAsync(c->D.cold + m * n, h.data() + m * n, n * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream), "batch cold start");
    }
    return batch_upload_consts(c);
}

int tdt_batch_set_state(tdt_ctx *c, int32_t icase, const double *cst, double rinit) {
    CHECK_CTX(c);
    int first = 0, cnt = 0;
    int rc = batch_case_range(c, icase, &first, &cnt);
    if (rc) return rc;
    if (!cst) return TDT_ERR_ARG;
    const tdt::BatchInfo &bi = *c->batch;
    const size_t nt = c->D.nt, n = c->pm.n;
    std::vector<double> h(5 * (size_t)cnt);
    for (int i = 0; i < cnt; ++i) {
        const int old = c->pm.cell_new2old[first + i] - bi.orig_cell_off[icase];
        for (int m = 0; m < 5; ++m) h[(size_t)m * cnt + i] = cst[(size_t)m * cnt + old];
    }
    for (int m = 0; m < 5; ++m) {
        HIPCALL(c, hipMemcpyAsync(c->D.c + m * nt + first, h.data() + (size_t)m * cnt,
                                  cnt * sizeof(double), hipMemcpyHostToDevice, c->stream),
                "batch state upload");
        HIPCALL(c, hipMemcpyAsync(c->D.cold + m * n + first, h.data() + (size_t)m * cnt,
                                  cnt * sizeof(double), hipMemcpyHostToDevice, c->stream),
                "batch state upload");
    }
    c->batch->rinit[icase] = rinit;
    return batch_upload_consts(c);
}

int tdt_batch_get_state(tdt_ctx *c, int32_t icase, double *cst) {
    CHECK_CTX(c);
    int first = 0, cnt = 0;
    int rc = batch_case_range(c, icase, &first, &cnt);
    if (rc) return rc;
    if (!cst) return TDT_ERR_ARG;
    const size_t nt = c->D.nt;
    std::vector<double> h(5 * (size_t)cnt);
    for (int m = 0; m < 5; ++m)
        HIPCALL(c, hipMemcpyAsync(h.data() + (size_t)m * cnt, c->D.c + m * nt + first,
                                  cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                "batch state download");
    rc = sync(c);
    if (rc) return rc;
    const tdt::BatchInfo &bi = *c->batch;
    for (int i = 0; i < cnt; ++i) {
        const int old = c->pm.cell_new2old[first + i] - bi.orig_cell_off[icase];
        for (int m = 0; m < 5; ++m) cst[(size_t)m * cnt + old] = h[(size_t)m * cnt + i];
    }
    return TDT_OK;
}

int tdt_batch_iterate(tdt_ctx *c, int32_t niter, double *l2, double *linf,
                      int32_t *first_nonfinite) {
    CHECK_CTX(c);
    if (!c->batch) {
        c->err = "not a batch context";
        return TDT_ERR_ARG;
    }
    if (niter < 0) return TDT_ERR_ARG;
    if (first_nonfinite) *first_nonfinite = 0;
    if (niter == 0) return TDT_OK;
    return tdt::fused_iterate(c, niter, l2, linf, first_nonfinite);
}

}  // extern "C"
