// C ABI of libtdt (include/tdt.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "api_internal.hpp"
#include "reduce.hpp"

using tdt::Consts;
using tdt::Dev;
using namespace tdt::api;

namespace tdt {
namespace api {

thread_local std::string g_create_error;

}  // namespace api

void api_set_error(const std::string &msg) { api::g_create_error = msg; }

namespace api {

void set_rinit(tdt_ctx *c, double rinit) {
    c->rinit = rinit;
    double rtol = rinit * c->params_rtred;           // RTOLERANCE = RINIT*RTRED
    if (rtol != c->K.rtol) tdt::fused_drop_graphs(c);
    c->K.rtol = rtol;
    c->K.rtol_rtred = rtol * c->params_rtred;        // RTOLERANCE*RTRED (viscous.f90:120)
    c->Kf.rtol = (float)rtol;
    c->Kf.rtol_rtred = (float)c->K.rtol_rtred;
}

}  // namespace api
}  // namespace tdt

namespace {

// ---- named arrays --------------------------------------------------------------
enum Kind { CELLGHOST, CELL, NODE, SCALAR };
struct ArrRef {
    double *p;
    Kind kind;
};

bool lookup(tdt_ctx *c, const char *name, ArrRef &a) {
    Dev &D = c->D;
    const size_t nt = D.nt, n = D.n, m = D.m;
    struct Fam { const char *pre; double *base; size_t stride; Kind kind; };
    const Fam fams[] = {
        {"COLD", D.cold, n, CELL}, {"C", D.c, nt, CELLGHOST}, {"F", D.f, nt, CELLGHOST},
        {"G", D.gf, nt, CELLGHOST}, {"R", D.r, n, CELL}, {"D2Q", D.d2q, nt, CELLGHOST},
        {"D24", D.d24, n, CELL},
    };
    for (const Fam &f : fams) {
        size_t l = std::strlen(f.pre);
        if (!std::strncmp(name, f.pre, l) && std::strlen(name) == l + 1 && name[l] >= '1' &&
            name[l] <= '5') {
            a.p = f.base + (size_t)(name[l] - '1') * f.stride;
            a.kind = f.kind;
            return true;
        }
    }
    struct One { const char *nm; double *p; Kind kind; };
    const One ones[] = {
        {"U", D.u, CELLGHOST}, {"V", D.v, CELLGHOST}, {"T", D.t, CELLGHOST},
        {"P", D.p, CELLGHOST}, {"DT", D.dt, CELLGHOST}, {"ANU", D.anu, CELLGHOST},
        {"ANURT", D.anurt, CELLGHOST}, {"XCC", D.xcc, CELLGHOST}, {"YCC", D.ycc, CELLGHOST},
        {"VOL", D.vol, CELLGHOST}, {"DL", D.dl, CELL}, {"VISU", D.vis, CELL},
        {"VISV", D.vis + n, CELL}, {"VIST", D.vis + 2 * n, CELL}, {"VISR", D.vis + 3 * n, CELL},
        {"SOURCE3", D.src3, CELL}, {"SOURCE5", D.src5, CELL}, {"UN", D.nodeavg, NODE},
        {"VN", D.nodeavg + m, NODE}, {"TN", D.nodeavg + 2 * m, NODE},
        {"RHON", D.nodeavg + 3 * m, NODE}, {"RN", D.nodeavg + 4 * m, NODE},
        {"SC1X", D.sc, CELL}, {"SC1Y", D.sc + n, CELL}, {"SC2X", D.sc + 2 * n, CELL},
        {"SC2Y", D.sc + 3 * n, CELL}, {"SC3X", D.sc + 4 * n, CELL}, {"SC3Y", D.sc + 5 * n, CELL},
        {"SC4X", D.sc + 6 * n, CELL}, {"SC4Y", D.sc + 7 * n, CELL},
    };
    for (const One &o : ones)
        if (!std::strcmp(name, o.nm)) {
            a.p = o.p;
            a.kind = o.kind;
            return true;
        }
    return false;
}

int32_t kind_len(tdt_ctx *c, Kind k) {
    switch (k) {
    case CELLGHOST: return c->pm.nmax;
    case CELL: return c->pm.n;
    case NODE: return c->pm.m;
    default: return 1;
    }
}

int launched(tdt_ctx *c, hipError_t e, const char *what) {
    if (e != hipSuccess) return hip_fail(c, e, what);
    return TDT_OK;
}

}  // namespace

namespace tdt {
namespace api {

int sync(tdt_ctx *c) {
    HIPCALL(c, hipStreamSynchronize(c->stream), "stream synchronize");
    return TDT_OK;
}

bool params_ok(const tdt_params &P) {
    if (!(P.gama > 1.0) || !(P.amol > 0.0) || !(P.cfl > 0.0) || !(P.prandtl > 0.0) ||
        !(P.prandtlt > 0.0) || !(P.t0 > 0.0) || !(P.p0 > 0.0) || (P.iaxis != 0 && P.iaxis != 1)) {
        g_create_error = "flow parameters out of range (GAMA>1, AMOL, CFL, PRANDTL, PRANDTLT, P0, "
                         "T0 > 0, IAXIS in {0,1})";
        return false;
    }
    return true;
}

int init_device(tdt_ctx *c, const tdt_params *P) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_create_error = "no HIP device available";
        return TDT_ERR_DEVICE;
    }
    hipGetDevice(&c->device);
    c->params_rtred = P->rtred;
    c->params = *P;
    c->K = make_consts<double>(*P);
    c->Kf = make_consts<float>(*P);
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        g_create_error = "hipStreamCreate failed";
        return TDT_ERR_DEVICE;
    }
    const tdt::PreparedMesh &pm = c->pm;
    Dev &D = c->D;
    D.n = pm.n; D.g = pm.g; D.m = pm.m; D.nt = pm.n + pm.g;
    const size_t n = pm.n, nt = D.nt, m = pm.m, g = pm.g;
    bool ok = true;
    auto up = [&](void *dst, const void *src, size_t bytes) {
        if (ok && bytes) ok = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    int4 *nc4 = dev_alloc<int4>(c, n), *nod4 = dev_alloc<int4>(c, n);
    double *xn = dev_alloc<double>(c, m), *yn = dev_alloc<double>(c, m);
    int *gpar = dev_alloc<int>(c, g), *gside = dev_alloc<int>(c, g), *gty = dev_alloc<int>(c, g),
        *gopp = dev_alloc<int>(c, g);
    int *nptr = dev_alloc<int>(c, m + 1), *nlist = dev_alloc<int>(c, pm.nlist.size());
    int4 *novr = dev_alloc<int4>(c, m);
    int *g6ptr = dev_alloc<int>(c, g + 1), *g6cell = dev_alloc<int>(c, pm.g6cell.size());
    double *g6dely = dev_alloc<double>(c, pm.g6dely.size());
    ok = nc4 && nod4 && xn && yn && gpar && gside && gty && gopp && nptr && nlist && novr &&
         g6ptr && g6cell && g6dely;
    up(nc4, pm.nc4.data(), n * sizeof(int4));
    up(nod4, pm.nod4.data(), n * sizeof(int4));
    up(xn, pm.xn.data(), m * sizeof(double));
    up(yn, pm.yn.data(), m * sizeof(double));
    up(gpar, pm.gparent.data(), g * sizeof(int));
    up(gside, pm.gside.data(), g * sizeof(int));
    up(gty, pm.gtype.data(), g * sizeof(int));
    up(gopp, pm.gopp.data(), g * sizeof(int));
    up(nptr, pm.nptr.data(), (m + 1) * sizeof(int));
    up(nlist, pm.nlist.data(), pm.nlist.size() * sizeof(int));
    up(novr, pm.novr.data(), m * sizeof(int4));
    up(g6ptr, pm.g6ptr.data(), (g + 1) * sizeof(int));
    up(g6cell, pm.g6cell.data(), pm.g6cell.size() * sizeof(int));
    up(g6dely, pm.g6dely.data(), pm.g6dely.size() * sizeof(double));
    D.nc4 = nc4; D.nod4 = nod4; D.xn = xn; D.yn = yn;
    D.gparent = gpar; D.gside = gside; D.gtype = gty; D.gopp = gopp;
    D.nptr = nptr; D.nlist = nlist; D.novr = novr;
    D.g6ptr = g6ptr; D.g6cell = g6cell; D.g6dely = g6dely;
    if (ok) {
        D.xcc = dev_alloc<double>(c, nt); D.ycc = dev_alloc<double>(c, nt);
        D.vol = dev_alloc<double>(c, nt); D.dl = dev_alloc<double>(c, n);
        D.sc = dev_alloc<double>(c, 8 * n);
        D.c = dev_alloc<double>(c, 5 * nt); D.cold = dev_alloc<double>(c, 5 * n);
        D.u = dev_alloc<double>(c, nt); D.v = dev_alloc<double>(c, nt);
        D.t = dev_alloc<double>(c, nt); D.p = dev_alloc<double>(c, nt);
        D.dt = dev_alloc<double>(c, nt);
        D.f = dev_alloc<double>(c, 5 * nt); D.gf = dev_alloc<double>(c, 5 * nt);
        D.r = dev_alloc<double>(c, 5 * n); D.d2q = dev_alloc<double>(c, 5 * nt);
        D.d24 = dev_alloc<double>(c, 5 * n);
        D.anu = dev_alloc<double>(c, nt); D.anurt = dev_alloc<double>(c, nt);
        D.vis = dev_alloc<double>(c, 4 * n);
        D.src3 = dev_alloc<double>(c, n); D.src5 = dev_alloc<double>(c, n);
        D.nodeavg = dev_alloc<double>(c, 5 * m);
        D.scal = dev_alloc<double>(c, 16);
        D.part = dev_alloc<double>(c, (size_t)tdt::nblocks(pm.n) * tdt::kPartStride);
        D.flags = dev_alloc<int>(c, 4);
        ok = D.xcc && D.ycc && D.vol && D.dl && D.sc && D.c && D.cold && D.u && D.v && D.t &&
             D.p && D.dt && D.f && D.gf && D.r && D.d2q && D.d24 && D.anu && D.anurt && D.vis &&
             D.src3 && D.src5 && D.nodeavg && D.scal && D.part && D.flags;
    }
    if (ok) ok = tdt::launch_geometry(c, P->iaxis) == hipSuccess;
    if (ok) ok = hipStreamSynchronize(c->stream) == hipSuccess;
    if (!ok) {
        g_create_error = "device allocation, upload or geometry kernel failed";
        return TDT_ERR_DEVICE;
    }
    return TDT_OK;
}

void cold_state(const tdt_params &P, const Consts<double> &K, double q[5], double *rinit) {
    double gama = (double)K.gama;
    double am = P.amachinit;
    double dynratio = 1.0 + (gama - 1.0) * 0.5 * (am * am);
    double tinit = (double)K.t0 / dynratio;
    double pinit = P.p0 / std::pow(dynratio, gama / (gama - 1.0));
    double uinit = am * std::sqrt(gama * K.rr * tinit);
    double vinit = 0.0;
    double einit = K.cv * tinit + 0.5 * (uinit * uinit + vinit * vinit);
    double rhoinit = pinit / (K.rr * tinit);
    double amuinit = (0.00000011848) * std::sqrt(P.amol) * std::pow(tinit, 0.6);
    *rinit = amuinit / rhoinit;
    q[0] = rhoinit;
    q[1] = rhoinit * uinit;
    q[2] = rhoinit * vinit;
    q[3] = rhoinit * einit;
    q[4] = rhoinit * *rinit;
}

}  // namespace api
}  // namespace tdt

extern "C" {

const char *tdt_version(void) { return "tdt 0.1 (gfx950)"; }

const char *tdt_last_error(const tdt_ctx *c) {
    return c ? c->err.c_str() : g_create_error.c_str();
}

int tdt_create(tdt_ctx **out, const tdt_mesh *mesh, const tdt_params *P, int32_t precision,
               int32_t ordering) {
    g_create_error.clear();
    if (!out || !mesh || !P) {
        g_create_error = "null argument";
        return TDT_ERR_ARG;
    }
    *out = nullptr;
    if (precision != TDT_FP64 && precision != TDT_FP32_MIXED) {
        g_create_error = "unknown precision";
        return TDT_ERR_ARG;
    }
    if (!params_ok(*P)) return TDT_ERR_ARG;
    tdt_ctx *c = new tdt_ctx();
    c->precision = precision;
    int code = TDT_OK;
    std::string msg = tdt::prepare_mesh(*mesh, ordering, c->pm, &code);
    if (!msg.empty()) {
        g_create_error = msg;
        delete c;
        return code;
    }
    code = init_device(c, P);
    if (code != TDT_OK) {
        tdt_destroy(c);
        return code;
    }
    set_rinit(c, 0.0);
    *out = c;
    return TDT_OK;
}

int tdt_destroy(tdt_ctx *c) {
    if (!c) return TDT_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    tdt::fused_free(c);
    for (void *p : c->allocs) hipFree(p);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c->batch;
    delete c->adj;
    delete c->iadj;
    delete c;
    return TDT_OK;
}

int tdt_info(const tdt_ctx *c, int64_t *info) {
    if (!c || !info) return TDT_ERR_ARG;
    info[0] = c->pm.n; info[1] = c->pm.g; info[2] = c->pm.m; info[3] = c->pm.nmax;
    info[4] = c->pm.bandwidth; info[5] = c->precision; info[6] = c->pm.ordering_used;
    info[7] = c->device;
    return TDT_OK;
}

int tdt_cold_start(tdt_ctx *c, double *rinit_out) {
    CHECK_SINGLE(c);
    double q[5], rinit;
    cold_state(c->params, c->K, q, &rinit);
    set_rinit(c, rinit);
    HIPCALL(c, tdt::launch_cold_start(c, q[0], q[1], q[2], q[3], q[4]), "cold start");
    if (rinit_out) *rinit_out = rinit;
    return sync(c);
}

int tdt_set_state(tdt_ctx *c, const double *cst, double rinit) {
    CHECK_SINGLE(c);
    if (!cst) return TDT_ERR_ARG;
    const size_t n = c->pm.n, nt = c->D.nt;
    std::vector<double> h(5 * n);
    for (int q = 0; q < 5; ++q)
        for (size_t i = 0; i < n; ++i) h[q * n + i] = cst[q * n + c->pm.cell_new2old[i]];
    for (int q = 0; q < 5; ++q) {
        HIPCALL(c, hipMemcpyAsync(c->D.c + q * nt, h.data() + q * n, n * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream), "state upload");
        HIPCALL(c, hipMemcpyAsync(c->D.cold + q * n, h.data() + q * n, n * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream), "state upload");
    }
    set_rinit(c, rinit);
    return sync(c);
}

int tdt_get_state(tdt_ctx *c, double *cst) {
    CHECK_SINGLE(c);
    if (!cst) return TDT_ERR_ARG;
    const size_t n = c->pm.n, nt = c->D.nt;
    std::vector<double> h(5 * n);
    for (int q = 0; q < 5; ++q)
        HIPCALL(c, hipMemcpyAsync(h.data() + q * n, c->D.c + q * nt, n * sizeof(double),
                                  hipMemcpyDeviceToHost, c->stream), "state download");
    int rc = sync(c);
    if (rc) return rc;
    for (int q = 0; q < 5; ++q)
        for (size_t i = 0; i < n; ++i) cst[q * n + c->pm.cell_new2old[i]] = h[q * n + i];
    return TDT_OK;
}

int tdt_timestep(tdt_ctx *c) { CHECK_SINGLE(c); return launched(c, tdt::launch_timestep(c), "timestep"); }
int tdt_dissipation(tdt_ctx *c) { CHECK_SINGLE(c); return launched(c, tdt::launch_dissipation(c), "dissipation"); }
int tdt_primitives(tdt_ctx *c) { CHECK_SINGLE(c); return launched(c, tdt::launch_primitives(c), "primitives"); }
int tdt_apply_boundary(tdt_ctx *c) { CHECK_SINGLE(c); return launched(c, tdt::launch_boundary(c), "apply_boundary"); }
int tdt_viscous(tdt_ctx *c) { CHECK_SINGLE(c); return launched(c, tdt::launch_viscous(c), "viscous"); }
int tdt_source_terms(tdt_ctx *c) { CHECK_SINGLE(c); return launched(c, tdt::launch_source(c), "source_terms"); }
int tdt_residual(tdt_ctx *c) { CHECK_SINGLE(c); return launched(c, tdt::launch_residual(c), "residual"); }
int tdt_update(tdt_ctx *c, double alpha) { CHECK_SINGLE(c); return launched(c, tdt::launch_update(c, alpha), "update"); }

int tdt_monitor(tdt_ctx *c, double *l2, double *linf) {
    CHECK_SINGLE(c);
    HIPCALL(c, hipMemsetAsync(c->D.flags, 0, sizeof(int), c->stream), "monitor");
    HIPCALL(c, tdt::launch_monitor(c, 1, c->D.scal + 2), "monitor");
    double h[10];
    int flag = 0;
    HIPCALL(c, hipMemcpyAsync(h, c->D.scal + 2, sizeof h, hipMemcpyDeviceToHost, c->stream), "monitor");
    HIPCALL(c, hipMemcpyAsync(&flag, c->D.flags, sizeof flag, hipMemcpyDeviceToHost, c->stream), "monitor");
    int rc = sync(c);
    if (rc) return rc;
    for (int q = 0; q < 5; ++q) {
        if (l2) l2[q] = h[q];
        if (linf) linf[q] = h[5 + q];
    }
    if (flag) {
        c->err = "non-finite value in the state";
        return TDT_ERR_NONFINITE;
    }
    return TDT_OK;
}

int tdt_postprocess(tdt_ctx *c, double *nodal) {
    CHECK_SINGLE(c);
    if (!nodal) return TDT_ERR_ARG;
    const tdt::PreparedMesh &pm = c->pm;
    if (!c->post && !(c->post = dev_alloc<double>(c, (size_t)6 * pm.m, false))) {
        c->err = "postprocess: device allocation failed";
        return TDT_ERR_DEVICE;
    }
    HIPCALL(c, tdt::launch_postprocess(c, c->post), "postprocess");
    std::vector<double> h((size_t)6 * pm.m);
    HIPCALL(c, hipMemcpyAsync(h.data(), c->post, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream), "postprocess");
    int rc = sync(c);
    if (rc) return rc;
    for (int k = 0; k < 6; ++k)
        for (int q = 0; q < pm.m; ++q)
            nodal[(size_t)k * pm.m + pm.node_new2old[q]] = h[(size_t)k * pm.m + q];
    return TDT_OK;
}

int tdt_iterate(tdt_ctx *c, int32_t niter, double *l2, double *linf, int32_t *first_nonfinite) {
    CHECK_SINGLE(c);
    if (niter < 0) return TDT_ERR_ARG;
    if (first_nonfinite) *first_nonfinite = 0;
    if (niter == 0) return TDT_OK;
    return tdt::fused_iterate(c, niter, l2, linf, first_nonfinite);
}

int tdt_get_array(tdt_ctx *c, const char *name, double *out, int32_t len) {
    CHECK_CTX(c);
    if (!name || !out) return TDT_ERR_ARG;
    if (!std::strcmp(name, "RINIT") || !std::strcmp(name, "RTOL") || !std::strcmp(name, "DTMIN")) {
        if (len != 1) { c->err = "scalar length must be 1"; return TDT_ERR_ARG; }
        if (!std::strcmp(name, "RINIT")) { *out = c->rinit; return TDT_OK; }
        if (!std::strcmp(name, "RTOL")) { *out = c->K.rtol; return TDT_OK; }
        double v = 0;
        HIPCALL(c, hipMemcpyAsync(&v, c->D.scal, sizeof v, hipMemcpyDeviceToHost, c->stream), "get");
        int rc = sync(c);
        *out = v;
        return rc;
    }
    ArrRef a;
    if (!lookup(c, name, a)) { c->err = std::string("unknown array ") + name; return TDT_ERR_ARG; }
    int32_t want = kind_len(c, a.kind);
    if (len != want) {
        c->err = std::string("array ") + name + " has length " + std::to_string(want);
        return TDT_ERR_ARG;
    }
    const tdt::PreparedMesh &pm = c->pm;
    size_t cnt = a.kind == CELLGHOST ? (size_t)c->D.nt : a.kind == CELL ? pm.n : pm.m;
    std::vector<double> h(cnt);
    HIPCALL(c, hipMemcpyAsync(h.data(), a.p, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream), "get");
    int rc = sync(c);
    if (rc) return rc;
    if (a.kind == NODE) {
        for (int q = 0; q < pm.m; ++q) out[pm.node_new2old[q]] = h[q];
        return TDT_OK;
    }
    if (a.kind == CELLGHOST) std::memset(out, 0, sizeof(double) * (size_t)len);
    for (int q = 0; q < pm.n; ++q) out[pm.cell_new2old[q]] = h[q];
    if (a.kind == CELLGHOST)
        for (int k = 0; k < pm.g; ++k) out[pm.ghost_orig_id[k] - 1] = h[pm.n + k];
    return TDT_OK;
}

int tdt_set_array(tdt_ctx *c, const char *name, const double *in, int32_t len) {
    CHECK_CTX(c);
    if (!name || !in) return TDT_ERR_ARG;
    ArrRef a;
    if (!lookup(c, name, a)) { c->err = std::string("unknown array ") + name; return TDT_ERR_ARG; }
    int32_t want = kind_len(c, a.kind);
    if (len != want) {
        c->err = std::string("array ") + name + " has length " + std::to_string(want);
        return TDT_ERR_ARG;
    }
    const tdt::PreparedMesh &pm = c->pm;
    size_t cnt = a.kind == CELLGHOST ? (size_t)c->D.nt : a.kind == CELL ? pm.n : pm.m;
    std::vector<double> h(cnt);
    if (a.kind == NODE) {
        for (int q = 0; q < pm.m; ++q) h[q] = in[pm.node_new2old[q]];
    } else {
        for (int q = 0; q < pm.n; ++q) h[q] = in[pm.cell_new2old[q]];
        if (a.kind == CELLGHOST)
            for (int k = 0; k < pm.g; ++k) h[pm.n + k] = in[pm.ghost_orig_id[k] - 1];
    }
    HIPCALL(c, hipMemcpyAsync(a.p, h.data(), cnt * sizeof(double), hipMemcpyHostToDevice, c->stream), "set");
    const std::string nm(name);
    if (nm == "XCC" || nm == "YCC" || nm == "VOL" || nm == "DL" || nm.rfind("SC", 0) == 0)
        ++c->geom_version;   // the fused path refreshes its copies
    if (nm.rfind("SC", 0) == 0) c->sc_from_nodes = false;   // no longer the nodes' vectors
    return sync(c);
}

int tdt_set_profiling(tdt_ctx *c, int32_t enable) {
    CHECK_CTX(c);
    c->profiling = enable ? 1 : 0;
    return TDT_OK;
}

int tdt_kernel_times(tdt_ctx *c, double *times_ms) {
    CHECK_CTX(c);
    if (!times_ms) return TDT_ERR_ARG;
    for (int q = 0; q < tdt::kNumFusedKernels; ++q) times_ms[q] = c->ktimes[q];
    return TDT_OK;
}

int tdt_kernel_set(tdt_ctx *c, int32_t *set) {
    CHECK_CTX(c);
    if (!set) return TDT_ERR_ARG;
    *set = tdt::fused_kernel_set(c);
    return TDT_OK;
}

int tdt_synchronize(tdt_ctx *c) {
    CHECK_CTX(c);
    return sync(c);
}

}  // extern "C"
