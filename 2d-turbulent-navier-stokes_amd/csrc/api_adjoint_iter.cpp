// C ABI of libtdt, the adjoint of whole iterations: tdt_operators_eval,
// tdt_iteration_vjp, tdt_adjoint_forward and tdt_adjoint_reverse.
//
// The reverse of one iteration Q^n -> Q^{n+1} (turbnstwod.f90:95-145) given
// mu = dJ/dQ^{n+1}:
//   1. recompute the iteration from Q^n with the routine kernels, keeping the
//      stage states Q^(1..3), the residuals R(Q^(0..3)), D, VIS and DT (the
//      RK stages are recomputed, not stored);
//   2. for k = 4..1: the stage update's reverse (k_stage_vjp) and the residual
//      adjoint at Q^(k-1) (adjoint.hip) carry mu back through stage k;
//   3. the operator adjoint (adjoint_iter.hip) adds the paths through D, VIS
//      and DT, which read Q^n.
// An unrolled run keeps Q every `stride` ~ sqrt(N) iterations and recomputes
// one segment of states at a time in the reverse sweep: about 2 sqrt(N)
// states on the device for one extra forward pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "api_internal.hpp"

using namespace tdt::api;

namespace {

constexpr double kAlphas[4] = {0.25, 1.0 / 3.0, 0.5, 1.0};   // constants.f90:30-33

int iadj_setup(tdt_ctx *c) {
    if (c->iadj) return TDT_OK;
    if (c->params.local != 1) {
        c->err = "iteration adjoint: needs LOCAL = 1 (a global DTMIN is not differentiated)";
        return TDT_ERR_UNSUPPORTED;
    }
    int rc = adj_setup(c);   // residual adjoint buffers and node-slot lists
    if (rc) return rc;
    const tdt::PreparedMesh &pm = c->pm;
    const int n = pm.n, m = pm.m;
    const int *nc = pm.nc4.data();
    auto b = std::make_unique<tdt::IterAdjBufs>();
    // members of each cell: itself, then ascending the cells whose state its
    // D, VIS and DT read
    std::vector<int> mptr(n + 1, 0), mlist;
    mlist.reserve((size_t)14 * n);
    std::vector<int> set;
    for (int i = 0; i < n; ++i) {
        set.clear();
        auto add_cell = [&](int id) {
            if (id >= 0 && id < n && id != i) set.push_back(id);
        };
        auto add_ghost = [&](int gid) {   // a ghost's state reads its parent (and opposite)
            const int k = gid - n;
            add_cell(pm.gparent[k]);
            if (pm.gtype[k] == 3) add_cell(pm.gopp[k]);
        };
        for (int f = 0; f < 4; ++f) {
            const int j = nc[4 * i + f];
            if (j < n) {
                add_cell(j);
                for (int g = 0; g < 4; ++g) add_cell(nc[4 * j + g]);
            } else {
                add_ghost(j);
            }
        }
        for (int k = 0; k < 4; ++k) {
            const int v = pm.nod4[(size_t)4 * i + k];
            for (int e = pm.nptr[v]; e < pm.nptr[v + 1]; ++e) {
                const int id = pm.nlist[e];
                if (id < n) {
                    add_cell(id);
                } else {
                    add_ghost(id);
                    // a type-7 ghost's centroid enters through the node slots: its
                    // parent must be the cell or a face neighbour
                    const int par = pm.gparent[id - n];
                    bool near = par == i;
                    for (int f = 0; f < 4; ++f) near = near || nc[4 * i + f] == par;
                    if (pm.gtype[id - n] == 7 && !near) b->xy_ok = false;
                }
            }
        }
        std::sort(set.begin(), set.end());
        set.erase(std::unique(set.begin(), set.end()), set.end());
        mlist.push_back(i);
        mlist.insert(mlist.end(), set.begin(), set.end());
        mptr[i + 1] = (int)mlist.size();
        b->kmax = std::max(b->kmax, mptr[i + 1] - mptr[i]);
    }
    // transposed: for each member cell k, the (cell i, entry e) whose member is k, by i
    std::vector<int> tptr(n + 1, 0);
    for (int e = 0; e < (int)mlist.size(); ++e) ++tptr[mlist[e] + 1];
    for (int k = 0; k < n; ++k) tptr[k + 1] += tptr[k];
    std::vector<int2> tlist(mlist.size());
    std::vector<int> fill(tptr.begin(), tptr.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int e = mptr[i]; e < mptr[i + 1]; ++e) tlist[fill[mlist[e]]++] = make_int2(i, e);
    const size_t nm = mlist.size(), n5 = (size_t)5 * n;
    b->mptr = dev_alloc<int>(c, n + 1);
    b->mlist = dev_alloc<int>(c, nm);
    b->tptr = dev_alloc<int>(c, n + 1);
    b->tlist = dev_alloc<int2>(c, nm);
    b->JE = dev_alloc<double>(c, nm * 30, false);
    b->JEp = dev_alloc<double>(c, (size_t)15 * n, false);
    b->qn = dev_alloc<double>(c, n5);
    bool ok = b->mptr && b->mlist && b->tptr && b->tlist && b->JE && b->JEp && b->qn;
    for (int k = 0; k < 3; ++k) ok = ok && (b->qs[k] = dev_alloc<double>(c, n5));
    for (int k = 0; k < 4; ++k) ok = ok && (b->r[k] = dev_alloc<double>(c, n5));
    ok = ok && (b->mu = dev_alloc<double>(c, n5)) && (b->lam = dev_alloc<double>(c, n5)) &&
         (b->wr = dev_alloc<double>(c, n5)) && (b->W = dev_alloc<double>(c, n5)) &&
         (b->g = dev_alloc<double>(c, n5)) && (b->gdt = dev_alloc<double>(c, n)) &&
         (b->gvol = dev_alloc<double>(c, n)) && (b->grtol = dev_alloc<double>(c, n)) &&
         (b->part = dev_alloc<double>(c, (size_t)3 * tdt::nblocks(n))) &&
         (b->gp = dev_alloc<double>(c, 3)) && (b->gxy = dev_alloc<double>(c, (size_t)2 * m));
    ok = ok && hipMemcpy(b->mptr, mptr.data(), mptr.size() * sizeof(int),
                         hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(b->mlist, mlist.data(), nm * sizeof(int), hipMemcpyHostToDevice) ==
                   hipSuccess;
    ok = ok && hipMemcpy(b->tptr, tptr.data(), tptr.size() * sizeof(int),
                         hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(b->tlist, tlist.data(), nm * sizeof(int2), hipMemcpyHostToDevice) ==
                   hipSuccess;
    if (!ok) {
        c->err = "iteration adjoint: device allocation failed";
        return TDT_ERR_DEVICE;
    }
    c->iadj = b.release();
    return TDT_OK;
}

// node-derivative blocks, allocated when first asked for
int iadj_xy(tdt_ctx *c) {
    tdt::IterAdjBufs &b = *c->iadj;
    if (!b.xy_ok) {
        c->err = "iteration adjoint: node derivatives need every type-7 ghost in a cell's node "
                 "averages to belong to the cell or a face neighbour";
        return TDT_ERR_UNSUPPORTED;
    }
    if (!b.JEx) b.JEx = dev_alloc<double>(c, (size_t)280 * c->pm.n, false);
    if (!b.JEx) {
        c->err = "iteration adjoint: device allocation failed";
        return TDT_ERR_DEVICE;
    }
    return TDT_OK;
}

// One iteration with the routine kernels on the context's arrays, from the
// state in C (COLD <- C first).  With `keep`, the stage states Q^(1..3) and
// the residuals R(Q^(0..3)) go to the adjoint buffers.
hipError_t routine_iteration(tdt_ctx *c, bool keep) {
    tdt::IterAdjBufs &b = *c->iadj;
    hipError_t e;
    if ((e = tdt::launch_copy_cold(c)) != hipSuccess) return e;
    if ((e = tdt::launch_timestep(c)) != hipSuccess) return e;
    if ((e = tdt::launch_dissipation(c)) != hipSuccess) return e;
    for (int k = 0; k < 4; ++k) {
        if ((e = tdt::launch_primitives(c)) != hipSuccess) return e;
        if ((e = tdt::launch_boundary(c)) != hipSuccess) return e;
        if (k == 0 && (e = tdt::launch_viscous(c)) != hipSuccess) return e;
        if ((e = tdt::launch_source(c)) != hipSuccess) return e;
        if ((e = tdt::launch_residual(c)) != hipSuccess) return e;
        if (keep && (e = hipMemcpyAsync(b.r[k], c->D.r, (size_t)5 * c->D.n * sizeof(double),
                                        hipMemcpyDeviceToDevice, c->stream)) != hipSuccess)
            return e;
        if ((e = tdt::launch_update(c, kAlphas[k])) != hipSuccess) return e;
        if (keep && k < 3 && (e = tdt::launch_pack(c, c->D.c, b.qs[k])) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t zero(tdt_ctx *c, double *p, size_t count) {
    return hipMemsetAsync(p, 0, count * sizeof(double), c->stream);
}

// mu = dJ/dQ^{n+1} (b.mu) -> b.lam = dJ/dQ^n through one iteration from the
// state qn ([5][n], device); adds the explicit d/d(P0, T0, RTOL) to b.gp and,
// with `nodes`, d/d(x, y) (metres) to b.gxy.  Overwrites the context's C, COLD
// and routine arrays.
int iteration_vjp_core(tdt_ctx *c, const double *qn, bool nodes) {
    tdt::IterAdjBufs &b = *c->iadj;
    tdt::AdjBufs &a = *c->adj;
    const tdt::Dev &D = c->D;
    const int n = D.n;
    const size_t n5 = (size_t)5 * n;
    HIPCALL(c, tdt::launch_unpack(c, qn), "adjoint: state");
    HIPCALL(c, routine_iteration(c, true), "adjoint: forward recomputation");
    HIPCALL(c, zero(c, b.lam, n5), "adjoint: clear");
    HIPCALL(c, zero(c, b.W, n5), "adjoint: clear");
    HIPCALL(c, zero(c, b.gdt, n), "adjoint: clear");
    HIPCALL(c, zero(c, b.gvol, n), "adjoint: clear");
    HIPCALL(c, zero(c, b.grtol, n), "adjoint: clear");
    for (int k = 3; k >= 0; --k) {
        const double *qk = k == 0 ? qn : b.qs[k - 1];
        HIPCALL(c, tdt::launch_stage_vjp(c, kAlphas[k], qn, b.r[k], D.vis, D.d24, D.dt, D.vol,
                                         b.mu, b.lam, b.wr, b.W, b.gdt, b.gvol, b.grtol),
                "adjoint: stage update");
        HIPCALL(c, tdt::launch_res_vjp(c, qk, b.wr, a.tptr, a.tlist, a.J, a.Jp, a.part, a.g,
                                       a.gp, a.Jx, a.nptr, a.nlist, nodes ? a.gxy : nullptr),
                "adjoint: residual");
        HIPCALL(c, tdt::launch_add(c, 2, a.gp, b.gp), "adjoint: parameters");
        if (nodes) HIPCALL(c, tdt::launch_add(c, 2 * D.m, a.gxy, b.gxy), "adjoint: nodes");
        if (k > 0)
            HIPCALL(c, hipMemcpyAsync(b.mu, a.g, n5 * sizeof(double), hipMemcpyDeviceToDevice,
                                      c->stream), "adjoint: copy");
        else
            HIPCALL(c, tdt::launch_add(c, (int)n5, a.g, b.lam), "adjoint: accumulate");
    }
    HIPCALL(c, tdt::launch_op_jac(c, qn, b.mptr, b.mlist, b.kmax, b.JE, b.JEp,
                                   nodes ? b.JEx : nullptr),
            "adjoint: operator blocks");
    HIPCALL(c, tdt::launch_op_vjp(c, b.JE, b.JEp, nodes ? b.JEx : nullptr, b.W, b.gdt, b.gvol,
                                  b.grtol, b.tptr, b.tlist, a.nptr, a.nlist, b.part, b.g, b.gp,
                                  nodes ? b.gxy : nullptr),
            "adjoint: operators");
    HIPCALL(c, tdt::launch_add(c, (int)n5, b.g, b.lam), "adjoint: accumulate");
    return TDT_OK;
}

// Fixed-point adjoint at the state qn: the iteration recomputed once (stage
// states, residuals, D, VIS, DT kept in the buffers and routine arrays) and the
// residual blocks at the four stage states and the operator blocks formed once.
int fixed_point_prepare(tdt_ctx *c, const double *qn) {
    tdt::IterAdjBufs &b = *c->iadj;
    const int n = c->pm.n;
    for (int k = 0; k < 4; ++k) {
        if (!b.Jst[k]) b.Jst[k] = dev_alloc<double>(c, (size_t)125 * n, false);
        if (!b.Jpst[k]) b.Jpst[k] = dev_alloc<double>(c, (size_t)10 * n, false);
        if (!b.Jst[k] || !b.Jpst[k]) {
            c->err = "fixed-point adjoint: device allocation failed";
            return TDT_ERR_DEVICE;
        }
    }
    if (!b.gJ) b.gJ = dev_alloc<double>(c, (size_t)5 * n);
    if (!b.gpx) b.gpx = dev_alloc<double>(c, 3);
    if (!b.gJ || !b.gpx) {
        c->err = "fixed-point adjoint: device allocation failed";
        return TDT_ERR_DEVICE;
    }
    HIPCALL(c, tdt::launch_unpack(c, qn), "adjoint: state");
    HIPCALL(c, routine_iteration(c, true), "adjoint: forward recomputation");
    for (int k = 0; k < 4; ++k)
        HIPCALL(c, tdt::launch_res_jac(c, k == 0 ? qn : b.qs[k - 1], b.Jst[k], b.Jpst[k], nullptr),
                "adjoint: residual blocks");
    HIPCALL(c, tdt::launch_op_jac(c, qn, b.mptr, b.mlist, b.kmax, b.JE, b.JEp, nullptr),
            "adjoint: operator blocks");
    return TDT_OK;
}

// b.mu -> b.lam = Phi_Q^T mu + gj with the blocks fixed_point_prepare kept (no
// parameter or node derivatives): one fixed-point iteration
int frozen_vjp(tdt_ctx *c, const double *qn, const double *gj) {
    tdt::IterAdjBufs &b = *c->iadj;
    tdt::AdjBufs &a = *c->adj;
    const tdt::Dev &D = c->D;
    const int n = D.n;
    const size_t n5 = (size_t)5 * n;
    // cleared by a kernel: the fixed-point loop captures this in a graph
    HIPCALL(c, tdt::launch_zero_5n_5n_n(c, b.lam, b.W, b.gdt), "adjoint: clear");
    // b.gvol and b.grtol collect the stage reverses' volume and RTOL terms, which
    // nothing here reads (iteration_vjp_core clears them before it uses them);
    // the residual and operator reverses skip their parameter sums (gp nullptr)
    const double *mu = b.mu;   // stage k's adjoint input: b.mu, then the previous a.g
    for (int k = 3; k >= 0; --k) {
        HIPCALL(c, tdt::launch_stage_vjp(c, kAlphas[k], qn, b.r[k], D.vis, D.d24, D.dt, D.vol,
                                         mu, b.lam, b.wr, b.W, b.gdt, b.gvol, b.grtol),
                "adjoint: stage update");
        HIPCALL(c, tdt::launch_res_apply(c, b.Jst[k], b.Jpst[k], b.wr, a.tptr, a.tlist, a.part,
                                         a.g, nullptr), "adjoint: residual");
        if (k > 0)
            mu = a.g;   // the stream orders the next stage's read before a.g is rewritten
    }
    HIPCALL(c, tdt::launch_op_vjp(c, b.JE, b.JEp, nullptr, b.W, b.gdt, b.gvol, b.grtol, b.tptr,
                                  b.tlist, a.nptr, a.nlist, b.part, b.g, nullptr, nullptr),
            "adjoint: operators");
    // lam += stage 1's residual adjoint (a.g, which the operators leave alone), the
    // operators' (b.g) and gj, in that order
    HIPCALL(c, tdt::launch_add3(c, (int)n5, a.g, b.g, gj, b.lam), "adjoint: accumulate");
    return TDT_OK;
}

int download_params(tdt_ctx *c, double gp[3]) {
    HIPCALL(c, hipMemcpyAsync(gp, c->iadj->gp, 3 * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream), "download");
    return sync(c);
}

// b.gxy (metres, internal order) -> per grid.in millimetre, grid.in node order
int download_nodes(tdt_ctx *c, double *gxy) {
    const int m = c->pm.m;
    std::vector<double> h((size_t)2 * m);
    HIPCALL(c, hipMemcpyAsync(h.data(), c->iadj->gxy, h.size() * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream), "download");
    int rc = sync(c);
    if (rc) return rc;
    for (int v = 0; v < m; ++v) {
        const int o = c->pm.node_new2old[v];
        gxy[o] = h[v] * 0.001;
        gxy[(size_t)m + o] = h[(size_t)m + v] * 0.001;
    }
    return TDT_OK;
}

// d(C1..C5, RINIT)/d(P0, T0, AMACHINIT) of the cold-start state
// (turbnstwod.f90:53-71, the same expressions as cold_state)
void cold_state_derivs(const tdt_params &P, const tdt::Consts<double> &K, double dq[5][3],
                       double drinit[3]) {
    const double gama = K.gama, rr = K.rr, cv = K.cv, am = P.amachinit;
    const double dyn = 1.0 + (gama - 1.0) * 0.5 * (am * am);
    const double kex = gama / (gama - 1.0);
    const double t = (double)K.t0 / dyn;
    const double p = P.p0 / std::pow(dyn, kex);
    const double u = am * std::sqrt(gama * rr * t);
    const double e = cv * t + 0.5 * u * u;
    const double rho = p / (rr * t);
    const double mu = 0.00000011848 * std::sqrt(P.amol) * std::pow(t, 0.6);
    const double rinit = mu / rho;
    const double ddyn_dm = (gama - 1.0) * am;
    for (int k = 0; k < 3; ++k) {
        double dt = 0.0, dp = 0.0, du = 0.0;
        if (k == 0) dp = 1.0 / std::pow(dyn, kex);
        if (k == 1) dt = 1.0 / dyn;
        if (k == 2) {
            dt = -t / dyn * ddyn_dm;
            dp = -kex * p / dyn * ddyn_dm;
            du = std::sqrt(gama * rr * t);
        }
        du += u / (2.0 * t) * dt;
        const double de = cv * dt + u * du;
        const double drho = dp / (rr * t) - rho / t * dt;
        const double dmu = 0.6 * mu / t * dt;
        const double dri = dmu / rho - mu / (rho * rho) * drho;
        dq[0][k] = drho;
        dq[1][k] = drho * u + rho * du;
        dq[2][k] = 0.0;
        dq[3][k] = drho * e + rho * de;
        dq[4][k] = drho * rinit + rho * dri;
        drinit[k] = dri;
    }
}

// the context's C rows -> dst ([5][n])
int save_state(tdt_ctx *c, double *dst) {
    HIPCALL(c, tdt::launch_pack(c, c->D.c, dst), "adjoint: save state");
    return TDT_OK;
}

}  // namespace

extern "C" {

int tdt_operators_eval(tdt_ctx *c, const double *cst, double *out) {
    CHECK_SINGLE(c);
    if (!cst || !out) return TDT_ERR_ARG;
    int rc = iadj_setup(c);
    if (!rc) rc = upload_cells(c, cst, c->iadj->qn);
    if (rc) return rc;
    tdt::IterAdjBufs &b = *c->iadj;
    HIPCALL(c, tdt::launch_op_eval(c, b.qn, b.r[0], b.r[1]), "operators");
    rc = download_cells(c, b.r[0], out);
    if (!rc) rc = download_cells(c, b.r[1], out + (size_t)5 * c->pm.n);
    return rc;
}

int tdt_operators_vjp(tdt_ctx *c, const double *cst, double rinit, const double *w,
                      double *g, double *gparam, double *gxy) {
    CHECK_SINGLE(c);
    if (!cst || !w || !g) return TDT_ERR_ARG;
    int rc = iadj_setup(c);
    if (!rc && gxy) rc = iadj_xy(c);
    if (rc) return rc;
    tdt::IterAdjBufs &b = *c->iadj;
    const int n = c->pm.n;
    const double rinit0 = c->rinit;
    set_rinit(c, rinit);
    // weights: rows 0-4 (E) -> W, row 5 (DT) -> gdt, row 6 (VOL) -> gvol
    std::vector<double> h((size_t)7 * n);
    for (int r = 0; r < 7; ++r)
        for (int i = 0; i < n; ++i) h[(size_t)r * n + i] = w[(size_t)r * n + c->pm.cell_new2old[i]];
    rc = upload_cells(c, cst, b.qn);
    if (!rc) {
        HIPCALL(c, hipMemcpyAsync(b.W, h.data(), (size_t)5 * n * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream), "upload");
        HIPCALL(c, hipMemcpyAsync(b.gdt, h.data() + (size_t)5 * n, n * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream), "upload");
        HIPCALL(c, hipMemcpyAsync(b.gvol, h.data() + (size_t)6 * n, n * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream), "upload");
        HIPCALL(c, zero(c, b.grtol, n), "clear");
        HIPCALL(c, zero(c, b.gp, 3), "clear");
        if (gxy) HIPCALL(c, zero(c, b.gxy, (size_t)2 * c->pm.m), "clear");
        HIPCALL(c, tdt::launch_op_jac(c, b.qn, b.mptr, b.mlist, b.kmax, b.JE, b.JEp,
                                      gxy ? b.JEx : nullptr), "operator blocks");
        HIPCALL(c, tdt::launch_op_vjp(c, b.JE, b.JEp, gxy ? b.JEx : nullptr, b.W, b.gdt, b.gvol,
                                      b.grtol, b.tptr, b.tlist, c->adj->nptr, c->adj->nlist,
                                      b.part, b.g, b.gp, gxy ? b.gxy : nullptr),
                "operators");
        rc = download_cells(c, b.g, g);
    }
    double gp[3] = {0.0, 0.0, 0.0};
    if (!rc) rc = download_params(c, gp);
    if (!rc && gxy) rc = download_nodes(c, gxy);
    set_rinit(c, rinit0);
    if (rc) return rc;
    if (gparam) {
        gparam[0] = gp[0];
        gparam[1] = gp[1];
        gparam[2] = gp[2] * c->params_rtred;   // RTOL = RINIT * RTRED
    }
    return TDT_OK;
}

int tdt_iteration_vjp(tdt_ctx *c, const double *cst, double rinit, const double *mu,
                      double *lam, double *gparam, double *gxy) {
    CHECK_SINGLE(c);
    if (!cst || !mu || !lam) return TDT_ERR_ARG;
    int rc = iadj_setup(c);
    if (!rc && gxy) rc = iadj_xy(c);
    if (rc) return rc;
    tdt::IterAdjBufs &b = *c->iadj;
    const int n = c->pm.n;
    // the context's own state is put back afterwards
    std::vector<double> keep((size_t)5 * n);
    double *saved = b.g;
    HIPCALL(c, tdt::launch_pack(c, c->D.c, saved), "adjoint: save state");
    HIPCALL(c, hipMemcpyAsync(keep.data(), saved, keep.size() * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream), "adjoint: save state");
    const double rinit0 = c->rinit;
    set_rinit(c, rinit);
    rc = upload_cells(c, cst, b.qn);
    if (!rc) rc = upload_cells(c, mu, b.mu);
    if (!rc) HIPCALL(c, zero(c, b.gp, 3), "adjoint: clear");
    if (!rc && gxy) HIPCALL(c, zero(c, b.gxy, (size_t)2 * c->pm.m), "adjoint: clear");
    if (!rc) rc = iteration_vjp_core(c, b.qn, gxy != nullptr);
    if (!rc) rc = download_cells(c, b.lam, lam);
    double gp[3] = {0.0, 0.0, 0.0};
    if (!rc) rc = download_params(c, gp);
    if (!rc && gxy) rc = download_nodes(c, gxy);
    // restore C, COLD and RINIT
    set_rinit(c, rinit0);
    int rc2 = TDT_OK;
    if (hipMemcpyAsync(b.qn, keep.data(), keep.size() * sizeof(double), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess ||
        tdt::launch_unpack(c, b.qn) != hipSuccess)
        rc2 = TDT_ERR_DEVICE;
    if (!rc2) rc2 = sync(c);
    if (rc) return rc;
    if (rc2) return rc2;
    if (gparam) {
        gparam[0] = gp[0];
        gparam[1] = gp[1];
        gparam[2] = gp[2] * c->params_rtred;   // RTOL = RINIT * RTRED
    }
    return TDT_OK;
}

namespace {

// The iterations between two convergence checks of the fixed-point loop, captured as
// one graph: each is frozen_vjp (b.mu -> b.lam, dJ/dQ added), then the two buffers swap
// roles.  An even count leaves the roles where they started, so one graph per role
// assignment serves every span.  The loop only launches kernels (frozen_vjp clears
// its buffers with kernels, not hipMemsetAsync).
// Each graph is kept, like its executable, until the loop ends (as the fused path
// keeps its graphs).
struct SpanGraphs {
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};   // by which buffer holds mu
    ~SpanGraphs() {
        for (int r = 0; r < 2; ++r) {
            if (exec[r]) (void)hipGraphExecDestroy(exec[r]);
            if (graph[r]) (void)hipGraphDestroy(graph[r]);
        }
    }
};

int capture_span(tdt_ctx *c, const double *qn, int span, hipGraph_t *graph,
                 hipGraphExec_t *exec) {
    tdt::IterAdjBufs &b = *c->iadj;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        c->err = "fixed-point adjoint: graph capture failed";
        return TDT_ERR_DEVICE;
    }
    int rc = TDT_OK;
    for (int k = 0; k < span && !rc; ++k) {
        rc = frozen_vjp(c, qn, b.gJ);
        std::swap(b.mu, b.lam);
    }
    hipError_t e = hipStreamEndCapture(c->stream, graph);
    if (!rc && e == hipSuccess) e = hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0);
    if (!rc && e != hipSuccess) {
        c->err = std::string("fixed-point adjoint: graph: ") + hipGetErrorString(e);
        rc = TDT_ERR_DEVICE;
    }
    return rc;
}

}  // namespace

int tdt_adjoint_fixed_point(tdt_ctx *c, const double *cst, double rinit, const double *dj_dc,
                            int32_t max_iter, double tol, double *lam, double *gparam,
                            double *gxy, int32_t *iters, double *change) {
    CHECK_SINGLE(c);
    if (!cst || !dj_dc || !lam || max_iter < 1 || !(tol >= 0.0)) return TDT_ERR_ARG;
    int rc = iadj_setup(c);
    if (!rc && gxy) rc = iadj_xy(c);
    if (rc) return rc;
    tdt::IterAdjBufs &b = *c->iadj;
    const int n = c->pm.n;
    const size_t n5 = (size_t)5 * n;
    std::vector<double> keep(n5);
    HIPCALL(c, tdt::launch_pack(c, c->D.c, b.g), "adjoint: save state");
    HIPCALL(c, hipMemcpyAsync(keep.data(), b.g, n5 * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream), "adjoint: save state");
    const double rinit0 = c->rinit;
    set_rinit(c, rinit);
    rc = upload_cells(c, cst, b.qn);
    if (!rc) rc = upload_cells(c, dj_dc, b.mu);
    if (!rc) rc = fixed_point_prepare(c, b.qn);
    if (!rc) HIPCALL(c, hipMemcpyAsync(b.gJ, b.mu, n5 * sizeof(double), hipMemcpyDeviceToDevice,
                                       c->stream), "adjoint: copy");
    // lam_{k+1} = Phi_Q^T lam_k + dJ/dQ from lam_0 = dJ/dQ.  Every kCheck
    // iterations the host measures the one-step change d_k = max|lam_{k+1} -
    // lam_k| / max|lam_{k+1}|.  For a contraction of rate rho the remaining
    // error is about d_k rho / (1 - rho): rho is estimated from two successive
    // checks, rho^kCheck = d_k / d_{k-kCheck}, and the iteration stops when that
    // estimate is <= tol (d_k alone underestimates it when rho is near 1).
    // The kCheck - 1 iterations before each checked one run as a captured graph
    // (TDT_FIXED_POINT_EAGER set: launched one by one, the same kernels in the same
    // order, for comparison).
    constexpr int kCheck = 25, kSpan = kCheck - 1;
    static_assert(kSpan % 2 == 0, "a span must leave the buffers' roles unchanged");
    const bool eager = std::getenv("TDT_FIXED_POINT_EAGER") != nullptr;
    SpanGraphs spans;
    double *const mu_first = b.mu;
    std::vector<double> h0(n5), h1(n5);
    int it = 0;
    double rel = INFINITY, prev = INFINITY;
    while (!rc && it < max_iter && !(rel <= tol)) {
        if (!eager && it % kCheck == 0 && it + kCheck <= max_iter) {
            const int role = b.mu == mu_first ? 0 : 1;
            if (!spans.exec[role])
                rc = capture_span(c, b.qn, kSpan, &spans.graph[role], &spans.exec[role]);
            if (rc) break;
            if (hipGraphLaunch(spans.exec[role], c->stream) != hipSuccess) {
                c->err = "fixed-point adjoint: graph launch failed";
                rc = TDT_ERR_DEVICE;
                break;
            }
            it += kSpan;
        }
        const bool check = (it + 1) % kCheck == 0 || it + 1 == max_iter;
        if (check)   // lam_k
            HIPCALL(c, hipMemcpyAsync(h0.data(), b.mu, n5 * sizeof(double),
                                      hipMemcpyDeviceToHost, c->stream), "download");
        rc = frozen_vjp(c, b.qn, b.gJ);
        if (rc) break;
        ++it;
        if (check) {
            HIPCALL(c, hipMemcpyAsync(h1.data(), b.lam, n5 * sizeof(double),
                                      hipMemcpyDeviceToHost, c->stream), "download");
            rc = sync(c);
            if (rc) break;
            double dmax = 0.0, lmax = 0.0;
            for (size_t k = 0; k < n5; ++k) {
                dmax = std::max(dmax, std::fabs(h1[k] - h0[k]));
                lmax = std::max(lmax, std::fabs(h1[k]));
            }
            const double d = lmax > 0.0 ? dmax / lmax : 0.0;
            if (!std::isfinite(d)) {
                c->err = "fixed-point adjoint: the iteration diverged";
                rc = TDT_ERR_NONFINITE;
                break;
            }
            if (d == 0.0) {
                rel = 0.0;                                     // exactly converged
            } else if (std::isfinite(prev) && d < prev && it % kCheck == 0) {
                const double rho = std::pow(d / prev, 1.0 / kCheck);
                rel = rho < 1.0 ? d * rho / (1.0 - rho) : INFINITY;
            } else {
                rel = INFINITY;                                // not contracting yet
            }
            if (it == max_iter) rel = std::isfinite(rel) ? rel : d;
            prev = d;
        }
        std::swap(b.mu, b.lam);   // lam_{k+1} is the next iteration's input
    }
    // lam* out, then one full reverse iteration with mu = lam* for Phi_p^T lam*
    // and Phi_xy^T lam*
    if (!rc) rc = download_cells(c, b.mu, lam);
    if (!rc) HIPCALL(c, zero(c, b.gp, 3), "adjoint: clear");
    if (!rc && gxy) HIPCALL(c, zero(c, b.gxy, (size_t)2 * c->pm.m), "adjoint: clear");
    if (!rc) rc = iteration_vjp_core(c, b.qn, gxy != nullptr);
    double gp[3] = {0.0, 0.0, 0.0};
    if (!rc) rc = download_params(c, gp);
    if (!rc && gxy) rc = download_nodes(c, gxy);
    set_rinit(c, rinit0);
    int rc2 = TDT_OK;
    if (hipMemcpyAsync(b.qn, keep.data(), n5 * sizeof(double), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess ||
        tdt::launch_unpack(c, b.qn) != hipSuccess)
        rc2 = TDT_ERR_DEVICE;
    if (!rc2) rc2 = sync(c);
    if (iters) *iters = it;
    if (change) *change = rel;
    if (rc) return rc;
    if (rc2) return rc2;
    if (gparam) {
        gparam[0] = gp[0];
        gparam[1] = gp[1];
        gparam[2] = gp[2] * c->params_rtred;   // RTOL = RINIT * RTRED
    }
    return TDT_OK;   // *change > tol: max_iter reached first (the caller decides)
}

int tdt_adjoint_forward(tdt_ctx *c, int32_t niter, const double *c0, double rinit,
                        double *cn) {
    CHECK_SINGLE(c);
    if (niter < 1 || !cn) return TDT_ERR_ARG;
    int rc = iadj_setup(c);
    if (rc) return rc;
    tdt::IterAdjBufs &b = *c->iadj;
    const int n = c->pm.n;
    const size_t n5 = (size_t)5 * n;
    b.cold = c0 == nullptr;
    if (b.cold) {
        double q[5];
        cold_state(c->params, c->K, q, &rinit);
        HIPCALL(c, tdt::launch_cold_start(c, q[0], q[1], q[2], q[3], q[4]), "cold start");
    } else {
        rc = upload_cells(c, c0, b.qn);
        if (rc) return rc;
        HIPCALL(c, tdt::launch_unpack(c, b.qn), "adjoint: state");
    }
    set_rinit(c, rinit);
    b.rinit = rinit;
    const int stride = std::max(1, (int)std::ceil(std::sqrt((double)niter)));
    const int nck = (niter + stride - 1) / stride;
    while ((int)b.ck.size() < nck) {
        double *p = dev_alloc<double>(c, n5, false);
        if (!p) {
            c->err = "adjoint: checkpoint allocation failed";
            return TDT_ERR_DEVICE;
        }
        b.ck.push_back(p);
    }
    while ((int)b.seg.size() < stride) {
        double *p = dev_alloc<double>(c, n5, false);
        if (!p) {
            c->err = "adjoint: checkpoint allocation failed";
            return TDT_ERR_DEVICE;
        }
        b.seg.push_back(p);
    }
    if (!b.qlast && !(b.qlast = dev_alloc<double>(c, n5, false))) {
        c->err = "adjoint: allocation failed";
        return TDT_ERR_DEVICE;
    }
    b.niter = 0;
    b.stride = stride;
    for (int it = 0; it < niter; ++it) {
        if (it % stride == 0) HIPCALL(c, tdt::launch_pack(c, c->D.c, b.ck[it / stride]),
                                      "adjoint: checkpoint");
        HIPCALL(c, routine_iteration(c, false), "adjoint: forward");
    }
    HIPCALL(c, tdt::launch_copy_cold(c), "adjoint: forward");
    rc = save_state(c, b.qlast);
    if (!rc) rc = download_cells(c, b.qlast, cn);
    if (rc) return rc;
    b.niter = niter;
    return TDT_OK;
}

int tdt_adjoint_reverse(tdt_ctx *c, const double *dj_dcn, double *lam0, double *gparam,
                        double *gxy) {
    CHECK_SINGLE(c);
    if (!dj_dcn || !lam0) return TDT_ERR_ARG;
    if (!c->iadj || c->iadj->niter < 1) {
        c->err = "tdt_adjoint_reverse: no tdt_adjoint_forward run to reverse";
        return TDT_ERR_UNSUPPORTED;
    }
    int rc = gxy ? iadj_xy(c) : TDT_OK;
    if (rc) return rc;
    tdt::IterAdjBufs &b = *c->iadj;
    const int n = c->pm.n;
    const size_t n5 = (size_t)5 * n;
    set_rinit(c, b.rinit);
    rc = upload_cells(c, dj_dcn, b.mu);
    if (rc) return rc;
    HIPCALL(c, zero(c, b.gp, 3), "adjoint: clear");
    if (gxy) HIPCALL(c, zero(c, b.gxy, (size_t)2 * c->pm.m), "adjoint: clear");
    const int N = b.niter, s = b.stride;
    for (int seg = (N - 1) / s; seg >= 0; --seg) {
        const int first = seg * s, last = std::min(first + s, N);   // iterations [first, last)
        // recompute the segment's states Q^first .. Q^(last-1)
        HIPCALL(c, tdt::launch_unpack(c, b.ck[seg]), "adjoint: checkpoint");
        for (int it = first; it < last; ++it) {
            HIPCALL(c, tdt::launch_pack(c, c->D.c, b.seg[it - first]), "adjoint: segment");
            if (it + 1 < last) HIPCALL(c, routine_iteration(c, false), "adjoint: segment");
        }
        for (int it = last - 1; it >= first; --it) {
            rc = iteration_vjp_core(c, b.seg[it - first], gxy != nullptr);
            if (rc) return rc;
            HIPCALL(c, hipMemcpyAsync(b.mu, b.lam, n5 * sizeof(double), hipMemcpyDeviceToDevice,
                                      c->stream), "adjoint: copy");
        }
    }
    rc = download_cells(c, b.mu, lam0);
    double gp[3] = {0.0, 0.0, 0.0};
    if (!rc) rc = download_params(c, gp);
    if (!rc && gxy) rc = download_nodes(c, gxy);
    // the context's state goes back to the run's final state
    HIPCALL(c, tdt::launch_unpack(c, b.qlast), "adjoint: state");
    if (!rc) rc = sync(c);
    if (rc) return rc;
    if (gparam) {
        const double grinit = gp[2] * c->params_rtred;   // RTOL = RINIT * RTRED
        gparam[0] = gp[0];
        gparam[1] = gp[1];
        gparam[2] = 0.0;
        gparam[3] = grinit;
        if (b.cold) {
            // Q^0 and RINIT of a cold start depend on P0, T0 and AMACHINIT: sums
            // over the cells of lam0 (grid.in cell order) times dQ^0/dp
            double dq[5][3], dri[3];
            cold_state_derivs(c->params, c->K, dq, dri);
            double lsum[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            for (int q = 0; q < 5; ++q)
                for (int i = 0; i < n; ++i) lsum[q] += lam0[(size_t)q * n + i];
            for (int k = 0; k < 3; ++k) {
                double add = grinit * dri[k];
                for (int q = 0; q < 5; ++q) add += lsum[q] * dq[q][k];
                gparam[k] += add;
            }
        }
    }
    return TDT_OK;
}

}  // extern "C"
