// C ABI of libtdt (include/tdt.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "ctx.hpp"
#include "reduce.hpp"
#include "tdt.h"

using tdt::Consts;
using tdt::Dev;

namespace {

thread_local std::string g_create_error;

#define CHECK_CTX(c)                   \
    do {                               \
        if (!(c)) return TDT_ERR_ARG;  \
    } while (0)

// Entry points of one case: a batch context runs its cases through the
// tdt_batch_* calls only.
#define CHECK_SINGLE(c)                                                         \
    do {                                                                        \
        if (!(c)) return TDT_ERR_ARG;                                           \
        if ((c)->batch) {                                                       \
            (c)->err = std::string(__func__) + ": not available on a batch context"; \
            return TDT_ERR_UNSUPPORTED;                                         \
        }                                                                       \
    } while (0)

int hip_fail(tdt_ctx *c, hipError_t e, const char *what) {
    c->err = std::string(what) + ": " + hipGetErrorString(e);
    return TDT_ERR_DEVICE;
}

#define HIPCALL(c, call, what)                                  \
    do {                                                        \
        hipError_t e__ = (call);                                \
        if (e__ != hipSuccess) return hip_fail((c), e__, what); \
    } while (0)

template <class T>
T *dev_alloc(tdt_ctx *c, size_t count, bool zero = true) {
    void *p = nullptr;
    size_t bytes = count * sizeof(T) + 256;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    if (zero && hipMemset(p, 0, bytes) != hipSuccess) {
        hipFree(p);
        return nullptr;
    }
    c->allocs.push_back(p);
    return static_cast<T *>(p);
}

template <class R>
Consts<R> make_consts(const tdt_params &P) {
    // thermodynamics as in turbnstwod.f90:37-42; derived groupings keep the
    // reference's evaluation order so they are bit-identical to it
    double rr = 8314.0 / P.amol;
    double cv = rr / (P.gama - 1.0);
    double cp = cv * P.gama;
    double rho0 = P.p0 / (rr * P.t0);
    double s0 = P.p0 / std::pow(rho0, P.gama);
    double tinf = P.t0 / (1.0 + 0.5 * (P.gama - 1.0) * P.aminf * P.aminf);
    Consts<R> K{};
    K.gama = (R)P.gama; K.cfl = (R)P.cfl; K.prandtl = (R)P.prandtl; K.prandtlt = (R)P.prandtlt;
    K.t0 = (R)P.t0; K.pamb = (R)P.pamb; K.pinf = (R)P.pinf; K.deltabl = (R)P.deltabl;
    K.rr = (R)rr; K.cv = (R)cv; K.cp = (R)cp; K.s0 = (R)s0; K.tinf = (R)tinf;
    K.suth_sqrt_amol = (R)(tdt::kSuth * std::sqrt(P.amol));
    K.inv_1mg = (R)(1.0 / (1.0 - P.gama));
    K.sqcmu = (R)std::sqrt(2.0 * std::sqrt(tdt::kCmu));
    K.bl_u = (R)(2.0 * P.aminf * std::sqrt(P.gama * rr * tinf));
    K.gamma_rr = (R)(P.gama * rr);
    K.tflame = (R)P.tflame; K.ancomb = (R)P.ancomb; K.acomb = (R)P.acomb;
    K.rhoprop = (R)P.rhoprop;
    K.gth_08 = (R)std::pow(tdt::kGth, 0.8);
    K.local = P.local; K.npseudo = P.npseudo; K.iaxis = P.iaxis;
    K.pamb_set = P.pamb >= 0.0;
    return K;
}

void set_rinit(tdt_ctx *c, double rinit) {
    c->rinit = rinit;
    double rtol = rinit * c->params_rtred;           // RTOLERANCE = RINIT*RTRED
    c->K.rtol = rtol;
    c->K.rtol_rtred = rtol * c->params_rtred;        // RTOLERANCE*RTRED (viscous.f90:120)
    c->Kf.rtol = (float)rtol;
    c->Kf.rtol_rtred = (float)c->K.rtol_rtred;
}

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

int sync(tdt_ctx *c) {
    HIPCALL(c, hipStreamSynchronize(c->stream), "stream synchronize");
    return TDT_OK;
}

int launched(tdt_ctx *c, hipError_t e, const char *what) {
    if (e != hipSuccess) return hip_fail(c, e, what);
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

// Device side of a context whose mesh is prepared (c->pm): stream, uploads,
// work arrays and the geometry (geometry.f90:55-149).  On failure the caller
// destroys c.
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

// Uniform initial state of a cold start (turbnstwod.f90:53-71) with the host's
// libm, as the reference: q = C1..C5, *rinit = RINIT.
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

}  // namespace

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
