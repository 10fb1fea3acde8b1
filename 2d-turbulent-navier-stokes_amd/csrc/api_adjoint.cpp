// C ABI of libtdt, the residual and its adjoint (tdt_residual_eval,
// tdt_residual_vjp) and the parameter / node updates they are used with
// (tdt_set_params, tdt_set_nodes); kernels in adjoint.hip.
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "api_internal.hpp"

using namespace tdt::api;

namespace tdt {
namespace api {

int adj_setup(tdt_ctx *c) {
    if (c->adj) return TDT_OK;
    if (!c->pm.g6cell.empty()) {
        c->err = "residual adjoint: BC type 6 is not available";
        return TDT_ERR_UNSUPPORTED;
    }
    const tdt::PreparedMesh &pm = c->pm;
    const int n = pm.n;
    // transposed stencil: entries (i, s) with member s of cell i equal to k, by i
    std::vector<int> ptr(n + 1, 0);
    auto member = [&](int i, int s) { return s == 0 ? i : pm.nc4[(size_t)4 * i + s - 1]; };
    for (int i = 0; i < n; ++i)
        for (int s = 0; s < 5; ++s) {
            const int k = member(i, s);
            if (k < n) ++ptr[k + 1];
        }
    for (int k = 0; k < n; ++k) ptr[k + 1] += ptr[k];
    std::vector<int2> list(ptr[n]);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int s = 0; s < 5; ++s) {
            const int k = member(i, s);
            if (k < n) list[fill[k]++] = make_int2(i, s);
        }
    // node slots (the cell's nodes, then each face neighbour's): per-node lists of
    // the (cell, slot) that hold it, by cell
    const int m = pm.m;
    auto slot_node = [&](int i, int sl) {
        if (sl < 4) return pm.nod4[(size_t)4 * i + sl];
        const int j = pm.nc4[(size_t)4 * i + (sl - 4) / 4];
        return j < n ? pm.nod4[(size_t)4 * j + (sl - 4) % 4] : -1;
    };
    std::vector<int> nptr(m + 1, 0);
    for (int i = 0; i < n; ++i)
        for (int sl = 0; sl < 20; ++sl) {
            const int v = slot_node(i, sl);
            if (v >= 0) ++nptr[v + 1];
        }
    for (int v = 0; v < m; ++v) nptr[v + 1] += nptr[v];
    std::vector<int2> nlist(nptr[m]);
    std::vector<int> nfill(nptr.begin(), nptr.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int sl = 0; sl < 20; ++sl) {
            const int v = slot_node(i, sl);
            if (v >= 0) nlist[nfill[v]++] = make_int2(i, sl);
        }
    auto a = std::make_unique<tdt::AdjBufs>();
    a->Jx = dev_alloc<double>(c, (size_t)200 * n, false);
    a->gxy = dev_alloc<double>(c, (size_t)2 * m);
    a->nptr = dev_alloc<int>(c, m + 1);
    a->nlist = dev_alloc<int2>(c, nlist.size());
    a->q = dev_alloc<double>(c, (size_t)5 * n);
    a->v = dev_alloc<double>(c, (size_t)5 * n);
    a->g = dev_alloc<double>(c, (size_t)5 * n);
    a->r = dev_alloc<double>(c, (size_t)5 * n);
    a->J = dev_alloc<double>(c, (size_t)125 * n, false);
    a->Jp = dev_alloc<double>(c, (size_t)10 * n, false);
    a->part = dev_alloc<double>(c, (size_t)2 * tdt::nblocks(n));
    a->gp = dev_alloc<double>(c, 2);
    a->tptr = dev_alloc<int>(c, n + 1);
    a->tlist = dev_alloc<int2>(c, list.size());
    bool ok = a->q && a->v && a->g && a->r && a->J && a->Jp && a->part && a->gp && a->tptr &&
              a->tlist && a->Jx && a->gxy && a->nptr && a->nlist;
    ok = ok && hipMemcpy(a->nptr, nptr.data(), nptr.size() * sizeof(int),
                         hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(a->nlist, nlist.data(), nlist.size() * sizeof(int2),
                         hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(a->tptr, ptr.data(), ptr.size() * sizeof(int),
                         hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(a->tlist, list.data(), list.size() * sizeof(int2),
                         hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        c->err = "residual adjoint: device allocation failed";
        return TDT_ERR_DEVICE;
    }
    c->adj = a.release();
    return TDT_OK;
}

// [5][neles] in original ids <-> [5][n] internal order on the device
int upload_cells(tdt_ctx *c, const double *src, double *dst) {
    const size_t n = c->pm.n;
    std::vector<double> h(5 * n);
    for (int q = 0; q < 5; ++q)
        for (size_t i = 0; i < n; ++i) h[q * n + i] = src[q * n + c->pm.cell_new2old[i]];
    HIPCALL(c, hipMemcpyAsync(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                              c->stream), "upload");
    return sync(c);
}

int download_cells(tdt_ctx *c, const double *src, double *dst) {
    const size_t n = c->pm.n;
    std::vector<double> h(5 * n);
    HIPCALL(c, hipMemcpyAsync(h.data(), src, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream), "download");
    int rc = sync(c);
    if (rc) return rc;
    for (int q = 0; q < 5; ++q)
        for (size_t i = 0; i < n; ++i) dst[q * n + c->pm.cell_new2old[i]] = h[q * n + i];
    return TDT_OK;
}

}  // namespace api
}  // namespace tdt

extern "C" {

// ---- residual and its adjoint ------------------------------------------------------

int tdt_set_params(tdt_ctx *c, const tdt_params *P) {
    CHECK_SINGLE(c);
    if (!P) return TDT_ERR_ARG;
    if (!params_ok(*P)) {
        c->err = g_create_error;
        return TDT_ERR_ARG;
    }
    if (P->iaxis != c->params.iaxis) {
        c->err = "tdt_set_params: IAXIS fixes the geometry and cannot change";
        return TDT_ERR_UNSUPPORTED;
    }
    int rc = sync(c);
    if (rc) return rc;
    c->params = *P;
    c->params_rtred = P->rtred;
    c->K = make_consts<double>(*P);
    c->Kf = make_consts<float>(*P);
    set_rinit(c, c->rinit);
    tdt::fused_free(c);   // its graphs hold the old constants; rebuilt on the next iterate
    if (c->iadj) c->iadj->niter = 0;   // a kept forward run used the old constants
    return TDT_OK;
}

int tdt_set_nodes(tdt_ctx *c, const double *x_mm, const double *y_mm) {
    CHECK_SINGLE(c);
    if (!x_mm || !y_mm) return TDT_ERR_ARG;
    const tdt::PreparedMesh &pm = c->pm;
    std::vector<double> h((size_t)2 * pm.m);
    for (int v = 0; v < pm.m; ++v) {
        // X(N) = X(N) * MM_TO_M  (geometry.f90:35-36)
        h[v] = x_mm[pm.node_new2old[v]] * 0.001;
        h[(size_t)pm.m + v] = y_mm[pm.node_new2old[v]] * 0.001;
    }
    int rc = sync(c);
    if (rc) return rc;
    HIPCALL(c, hipMemcpyAsync(const_cast<double *>(c->D.xn), h.data(), pm.m * sizeof(double),
                              hipMemcpyHostToDevice, c->stream), "node upload");
    HIPCALL(c, hipMemcpyAsync(const_cast<double *>(c->D.yn), h.data() + pm.m,
                              pm.m * sizeof(double), hipMemcpyHostToDevice, c->stream),
            "node upload");
    HIPCALL(c, tdt::launch_geometry(c, c->params.iaxis), "geometry");
    ++c->geom_version;
    if (c->iadj) c->iadj->niter = 0;   // a kept forward run used the old geometry
    return sync(c);
}

int tdt_residual_eval(tdt_ctx *c, const double *cst, double *r) {
    CHECK_SINGLE(c);
    if (!cst || !r) return TDT_ERR_ARG;
    int rc = adj_setup(c);
    if (!rc) rc = upload_cells(c, cst, c->adj->q);
    if (rc) return rc;
    HIPCALL(c, tdt::launch_res_eval(c, c->adj->q, c->adj->r), "residual");
    return download_cells(c, c->adj->r, r);
}

int tdt_residual_vjp(tdt_ctx *c, const double *cst, const double *v, double *gq,
                     double *gparam, double *gxy) {
    CHECK_SINGLE(c);
    if (!cst || !v || !gq) return TDT_ERR_ARG;
    int rc = adj_setup(c);
    if (!rc) rc = upload_cells(c, cst, c->adj->q);
    if (!rc) rc = upload_cells(c, v, c->adj->v);
    if (rc) return rc;
    tdt::AdjBufs &a = *c->adj;
    HIPCALL(c, tdt::launch_res_vjp(c, a.q, a.v, a.tptr, a.tlist, a.J, a.Jp, a.part, a.g, a.gp,
                                   a.Jx, a.nptr, a.nlist, gxy ? a.gxy : nullptr),
            "residual adjoint");
    rc = download_cells(c, a.g, gq);
    if (rc) return rc;
    if (gparam) {
        double gp[2] = {0.0, 0.0};
        HIPCALL(c, hipMemcpyAsync(gp, a.gp, sizeof gp, hipMemcpyDeviceToHost, c->stream),
                "download");
        rc = sync(c);
        if (rc) return rc;
        gparam[0] = gp[0];
        gparam[1] = gp[1];
        gparam[2] = 0.0;   // R does not depend on AMACHINIT at a given state
    }
    if (gxy) {
        // per node, d/d(x, y) in metres -> grid.in millimetres (X = X_MM * 0.001)
        const int m = c->pm.m;
        std::vector<double> h((size_t)2 * m);
        HIPCALL(c, hipMemcpyAsync(h.data(), a.gxy, h.size() * sizeof(double),
                                  hipMemcpyDeviceToHost, c->stream), "download");
        rc = sync(c);
        if (rc) return rc;
        for (int v2 = 0; v2 < m; ++v2) {
            const int o = c->pm.node_new2old[v2];
            gxy[o] = h[v2] * 0.001;
            gxy[(size_t)m + o] = h[(size_t)m + v2] * 0.001;
        }
    }
    return TDT_OK;
}

}  // extern "C"
