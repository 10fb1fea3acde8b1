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
// compute them on the fly instead of running a separate boundary pass.
#include "ctx.hpp"
#include "reduce.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdlib>
#include <string>

namespace tdt {

struct FDev {
    int n, g, m, nt;
    int remap;                      // XCD-aware block order
    const int4 *nc4, *nod4;
    const double *xn, *yn;
    const int *gparent, *gtype, *gopp;
    const int *nptr, *nlist;
    const int4 *novr;
    const double *xcc, *ycc, *vol, *dl, *sc;
    double *qn;                     // [5][nt] Q^n (the C array), FP64
    void *st[2];                    // [5][n] stage states Q(1), Q(2)/Q(3)
    double *dt;                     // [n]
    void *anu, *anurt;              // [n]  (storage type SD)
    void *d2q;                      // [5][nt]
    void *dd;                       // [5][n] dissipation D
    void *vis;                      // [4][n]
    void *node;                     // [5][m] UN VN TN RHON RN
    unsigned long long *dtslot;     // DTMIN bits for LOCAL = 0
    double *part;                   // per-block monitor partials
    double *hist;                   // [niter][10]
    int *ctr;                       // 0: iteration, 1: ticket, 2: first non-finite
};

struct FusedBufs {
    FDev F{};
    int nb_cells = 0, nb_nodes = 0;
    size_t hist_cap = 0;
    std::vector<void *> own;
    hipGraph_t graph1 = nullptr, graphK = nullptr;
    hipGraphExec_t exec1 = nullptr, execK = nullptr;
    int chunk = 8;
    hipEvent_t ev[12];
    bool have_events = false;
    bool quad = false;               // face-parallel kernels (4 lanes per cell)
    int nbq_cells = 0, nbq_nodes = 0;
};

namespace {

// Logical block index so that each XCD (blocks are dispatched round-robin over
// the 8 XCDs) works on a contiguous range of cells and its L2 sees the halo.
__device__ __forceinline__ int logical_block(int remap) {
    int b = blockIdx.x, nb = gridDim.x;
    if (!remap || nb < 16) return b;
    int xcd = b & 7, local = b >> 3, per = nb >> 3, rem = nb & 7;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

template <class T>
__device__ __forceinline__ double ld(const void *p, size_t i) {
    return (double)static_cast<const T *>(p)[i];
}
template <class T>
__device__ __forceinline__ void stv(void *p, size_t i, double v) {
    static_cast<T *>(p)[i] = (T)v;
}

// Folds the pending stage-4 partials of the previous iteration into the
// history (one block; block 0 of k_pre and k_finalize call it).  The kernel
// boundary orders the partial writes before these reads.
__device__ __forceinline__ void fold_monitor(const FDev &F, int nblk) {
    if (F.ctr[3] == 0) return;   // uniform across the block
    __shared__ int flag;
    if (threadIdx.x == 0) flag = 0;
    __syncthreads();
    int it = F.ctr[0];
    reduce_partials(F.part, nblk, F.n, F.hist + (size_t)it * 10, &flag);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (flag && F.ctr[2] == 0) F.ctr[2] = it + 1;
        F.ctr[0] = it + 1;
        F.ctr[3] = 0;
    }
}

// ---------------------------------------------------------------------------------
// k_pre: cells [0, nb_cells) blocks, nodes after.
template <class SD, class R>
__global__ void __launch_bounds__(kBlock) k_pre(FDev F, Consts<double> K, Consts<R> KR,
                                                int nb_cells) {
    int b = blockIdx.x;
    if (b == 0) fold_monitor(F, nb_cells);
    if (b < nb_cells) {
        int i = b * kBlock + threadIdx.x;
        double dtl = 100000.0;
        if (i < F.n) {
            const int nt = F.nt;
            const double *Q = F.qn;
            double c0 = Q[i], c1 = Q[nt + i], c2 = Q[2 * nt + i], c3 = Q[3 * nt + i],
                   c4 = Q[4 * nt + i];
            double u, v, t, p;
            uvtp_from_q(c0, c1, c2, c3, u, v, t, p, K);
            double vel = sqrt(u * u + v * v);
            double dt = K.cfl * F.dl[i] / (fabs(vel) + sqrt(K.gamma_rr * t));
            if (K.local) F.dt[i] = dt;
            dtl = dt;
            int4 nb4 = F.nc4[i];
            int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
            // ghosts hold parent copies in DISSIPATION (solver_state.f90:77-88)
            double pn[4], cn[5][4];
            for (int k = 0; k < 4; ++k) {
                int j = nb[k] < F.n ? nb[k] : i;
                double q0 = Q[j], q1 = Q[nt + j], q2 = Q[2 * nt + j], q3 = Q[3 * nt + j];
                double uj, vj, tj, pj;
                uvtp_from_q(q0, q1, q2, q3, uj, vj, tj, pj, K);
                pn[k] = pj;
                cn[0][k] = q0; cn[1][k] = q1; cn[2][k] = q2; cn[3][k] = q3;
                cn[4][k] = Q[4 * nt + j];
            }
            double anun = fabs(pn[0] - p) + fabs(pn[1] - p) + fabs(pn[2] - p) + fabs(pn[3] - p);
            double anud = fabs(pn[0] + p) + fabs(pn[1] + p) + fabs(pn[2] + p) + fabs(pn[3] + p);
            stv<SD>(F.anu, i, anun / anud);
            anun = fabs(cn[4][0] - c4) + fabs(cn[4][1] - c4) + fabs(cn[4][2] - c4) +
                   fabs(cn[4][3] - c4);
            anud = fabs(cn[4][0] + c4) + fabs(cn[4][1] + c4) + fabs(cn[4][2] + c4) +
                   fabs(cn[4][3] + c4);
            stv<SD>(F.anurt, i, anun / dmax(anud, 8.0 * c0 * K.rtol));
            const double ci[5] = {c0, c1, c2, c3, c4};
            for (int m = 0; m < 5; ++m)
                stv<SD>(F.d2q, (size_t)m * nt + i,
                        cn[m][0] + cn[m][1] + cn[m][2] + cn[m][3] - 4.0 * ci[m]);
        }
        if (!K.local) block_atomic_min_pos(dtl, reinterpret_cast<double *>(F.dtslot));
        return;
    }
    int v = (b - nb_cells) * kBlock + threadIdx.x;
    if (v >= F.m) return;
    const int nt = F.nt;
    const double *Q = F.qn;
    int4 ov = F.novr[v];
    int j0 = F.nptr[v], j1 = F.nptr[v + 1];
    R au = R(0), av = R(0), at = R(0), arho = R(0), ar = ov.w ? KR.rtol_rtred : R(0);
    for (int j = j0; j < j1; ++j) {
        int id = F.nlist[j];
        R u, vv, t, rho, c5;
        if (id < F.n) {
            R p;
            rho = (R)Q[id];
            uvtp_from_q<R>(rho, (R)Q[nt + id], (R)Q[2 * nt + id], (R)Q[3 * nt + id], u, vv, t, p, KR);
            c5 = (R)Q[4 * nt + id];
        } else {
            int k = id - F.n, par = F.gparent[k], opp = F.gopp[k];
            Prim<R> pp, s;
            prim_from_q<R>(pp, (R)Q[par], (R)Q[nt + par], (R)Q[2 * nt + par], (R)Q[3 * nt + par],
                           (R)Q[4 * nt + par], KR);
            R uo = R(0), vo = R(0), to = R(0), po = R(0);
            int ty = F.gtype[k];
            if (ty == 3)
                uvtp_from_q<R>((R)Q[opp], (R)Q[nt + opp], (R)Q[2 * nt + opp], (R)Q[3 * nt + opp],
                               uo, vo, to, po, KR);
            ghost_state<R>(s, ty, pp, uo, vo, to, po, (R)F.ycc[id], KR);
            u = s.u; vv = s.v; t = s.t; rho = s.c[0]; c5 = s.c[4];
        }
        R rr = c5 / rho;
        if (j >= ov.x) au += u;
        if (j >= ov.y) av += vv;
        at += t;
        arho += rho;
        if (j >= ov.z) ar += rr;
    }
    R cnt = (R)(j1 - j0);
    const int m = F.m;
    stv<SD>(F.node, v, au / cnt);
    stv<SD>(F.node, (size_t)m + v, av / cnt);
    stv<SD>(F.node, (size_t)2 * m + v, at / cnt);
    stv<SD>(F.node, (size_t)3 * m + v, arho / cnt);
    stv<SD>(F.node, (size_t)4 * m + v, dmax(KR.rtol, ar / cnt));
}

// ---------------------------------------------------------------------------------
// k_dv: frozen dissipation D and viscous fluxes VIS from Q^n.
template <class SD, class R>
__global__ void __launch_bounds__(kBlock) k_dv(FDev F, Consts<double> K, Consts<R> KR) {
    int i = logical_block(F.remap) * kBlock + threadIdx.x;
    if (i >= F.n) return;
    const int nt = F.nt, n = F.n, m = F.m;
    const double *Q = F.qn;
    double dti = K.local ? F.dt[i] : __longlong_as_double((long long)*F.dtslot);
    if (!K.local) F.dt[i] = dti;
    int4 nb4 = F.nc4[i], nd4 = F.nod4[i];
    const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
    const int nd[4] = {nd4.x, nd4.y, nd4.z, nd4.w};
    // ---- dissipation (solver_state.f90:183-223); ghost neighbour = parent copy
    double ci[5];
    for (int q = 0; q < 5; ++q) ci[q] = Q[q * nt + i];
    R voli = (R)F.vol[i];
    R volbydt = voli / (R)dti;
    R anui = (R)ld<SD>(F.anu, i), anurti = (R)ld<SD>(F.anurt, i);
    R d2qi[5];
    for (int q = 0; q < 5; ++q) d2qi[q] = (R)ld<SD>(F.d2q, (size_t)q * nt + i);
    R d2[5], d4[5];
    for (int k = 0; k < 4; ++k) {
        int j = nb[k] < n ? nb[k] : i;
        R dtj = K.local ? (R)F.dt[j] : (R)dti;
        R volj = (R)F.vol[j];
        R a2, a4, b2, b4;
        jst_coeffs<R>(anui, (R)ld<SD>(F.anu, j), volbydt, volj, dtj, a2, a4);
        jst_coeffs<R>(anurti, (R)ld<SD>(F.anurt, j), volbydt, volj, dtj, b2, b4);
        for (int q = 0; q < 5; ++q) {
            R c2 = q < 4 ? a2 : b2, c4 = q < 4 ? a4 : b4;
            R dc = (R)(Q[q * nt + j] - ci[q]);
            R dq = (R)ld<SD>(F.d2q, (size_t)q * nt + j) - d2qi[q];
            if (k == 0) {
                d2[q] = c2 * dc;
                d4[q] = c4 * dq;
            } else {
                d2[q] = d2[q] + c2 * dc;
                d4[q] = d4[q] + c4 * dq;
            }
        }
    }
    for (int q = 0; q < 5; ++q) stv<SD>(F.dd, (size_t)q * n + i, d2[q] - d4[q]);
    // ---- viscous fluxes (viscous.f90:62-70, 130-217)
    R ua, va, ta, pa;
    R rhoa = (R)ci[0];
    uvtp_from_q<R>(rhoa, (R)ci[1], (R)ci[2], (R)ci[3], ua, va, ta, pa, KR);
    R ra = (R)ci[4] / rhoa;
    R xa = (R)F.xcc[i], ya = (R)F.ycc[i];
    Prim<R> pi;
    bool have_pi = false;
    R vu = R(0), vv = R(0), vt = R(0), vr = R(0);
    for (int f = 0; f < 4; ++f) {
        int bnode = nd[f], dnode = nd[(f + 1) & 3], c = nb[f];
        R uc, vc, tc, rhoc, c5c;
        if (c < n) {
            R pc;
            rhoc = (R)Q[c];
            uvtp_from_q<R>(rhoc, (R)Q[nt + c], (R)Q[2 * nt + c], (R)Q[3 * nt + c], uc, vc, tc, pc, KR);
            c5c = (R)Q[4 * nt + c];
        } else {
            if (!have_pi) {
                prim_from_q<R>(pi, rhoa, (R)ci[1], (R)ci[2], (R)ci[3], (R)ci[4], KR);
                have_pi = true;
            }
            int k = c - n, ty = F.gtype[k];
            R uo = R(0), vo = R(0), to = R(0), po = R(0);
            if (ty == 3) {
                int o = F.gopp[k];
                uvtp_from_q<R>((R)Q[o], (R)Q[nt + o], (R)Q[2 * nt + o], (R)Q[3 * nt + o], uo, vo,
                               to, po, KR);
            }
            Prim<R> s;
            ghost_state<R>(s, ty, pi, uo, vo, to, po, (R)F.ycc[c], KR);
            uc = s.u; vc = s.v; tc = s.t; rhoc = s.c[0]; c5c = s.c[4];
        }
        R rc = c5c / rhoc;
        ViscFace<R> o = viscous_face<R>(
            (R)F.sc[(2 * f) * n + i], (R)F.sc[(2 * f + 1) * n + i], xa, ya, (R)F.xn[bnode],
            (R)F.yn[bnode], (R)F.xcc[c], (R)F.ycc[c], (R)F.xn[dnode], (R)F.yn[dnode], ua, va, ta,
            rhoa, ra, (R)ld<SD>(F.node, bnode), (R)ld<SD>(F.node, (size_t)m + bnode),
            (R)ld<SD>(F.node, (size_t)2 * m + bnode), (R)ld<SD>(F.node, (size_t)3 * m + bnode),
            (R)ld<SD>(F.node, (size_t)4 * m + bnode), uc, vc, tc, rhoc, rc,
            (R)ld<SD>(F.node, dnode), (R)ld<SD>(F.node, (size_t)m + dnode),
            (R)ld<SD>(F.node, (size_t)2 * m + dnode), (R)ld<SD>(F.node, (size_t)3 * m + dnode),
            (R)ld<SD>(F.node, (size_t)4 * m + dnode), KR);
        vu = vu + o.u0 + o.u1;
        vv = vv + o.v0 + o.v1;
        vt = vt + o.t0 + o.t1;
        vr = vr + o.r0 + o.r1;
    }
    stv<SD>(F.vis, i, vu);
    stv<SD>(F.vis, (size_t)n + i, vv);
    stv<SD>(F.vis, (size_t)2 * n + i, vt);
    stv<SD>(F.vis, (size_t)3 * n + i, vr);
}

// ---------------------------------------------------------------------------------
// k_stage: one RK stage.  SI / SO: stage input / output storage types.
template <class R, class SI, class SO, class SD, int STAGE>
__global__ void __launch_bounds__(kBlock) k_stage(FDev F, Consts<double> K, Consts<R> KR,
                                                  double alpha) {
    int i = logical_block(F.remap) * kBlock + threadIdx.x;
    const int nt = F.nt, n = F.n;
    const void *qin = STAGE == 1 ? (const void *)F.qn : (const void *)F.st[STAGE & 1];
    const size_t sin = STAGE == 1 ? (size_t)nt : (size_t)n;
    double cnew[5] = {0, 0, 0, 0, 0}, cold[5] = {0, 0, 0, 0, 0};
    bool active = i < n;
    if (active) {
        Prim<R> pi;
        prim_from_q<R>(pi, (R)ld<SI>(qin, i), (R)ld<SI>(qin, sin + i), (R)ld<SI>(qin, 2 * sin + i),
                       (R)ld<SI>(qin, 3 * sin + i), (R)ld<SI>(qin, 4 * sin + i), KR);
        int4 nb4 = F.nc4[i];
        const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
        R acc[5];
        R gx[4], gy[4], gu[4], gv[4], gr[4];
        for (int f = 0; f < 4; ++f) {
            int j = nb[f];
            Prim<R> pj;
            if (j < n) {
                prim_from_q<R>(pj, (R)ld<SI>(qin, j), (R)ld<SI>(qin, sin + j),
                               (R)ld<SI>(qin, 2 * sin + j), (R)ld<SI>(qin, 3 * sin + j),
                               (R)ld<SI>(qin, 4 * sin + j), KR);
            } else {
                int k = j - n, ty = F.gtype[k];
                R uo = R(0), vo = R(0), to = R(0), po = R(0);
                if (ty == 3) {
                    int o = F.gopp[k];
                    uvtp_from_q<R>((R)ld<SI>(qin, o), (R)ld<SI>(qin, sin + o),
                                   (R)ld<SI>(qin, 2 * sin + o), (R)ld<SI>(qin, 3 * sin + o), uo,
                                   vo, to, po, KR);
                }
                ghost_state<R>(pj, ty, pi, uo, vo, to, po, (R)F.ycc[j], KR);
            }
            R sx = (R)F.sc[(2 * f) * n + i], sy = (R)F.sc[(2 * f + 1) * n + i];
            for (int q = 0; q < 5; ++q) {
                R tf = (pi.f[q] + pj.f[q]) * sx;
                R tg = (pi.g[q] + pj.g[q]) * sy;
                acc[q] = f == 0 ? tf + tg : acc[q] + tf + tg;
            }
            gx[f] = (R)F.xcc[j];
            gy[f] = (R)F.ycc[j];
            gu[f] = pj.u;
            gv[f] = pj.v;
            gr[f] = pj.c[4] / pj.c[0];
        }
        R s3;
        R ycci = (R)F.ycc[i];
        R s5 = source_terms<R>(gx[0], gx[1], gx[2], gx[3], gy[0], gy[1], gy[2], gy[3], gu[0], gu[1],
                               gu[2], gu[3], gv[0], gv[1], gv[2], gv[3], gr[0], gr[1], gr[2], gr[3],
                               pi.c[0], pi.c[4], pi.t, pi.p, pi.v, ycci, s3, KR);
        // residual.f90:34-60 and the update residual.f90:76-81 (FP64)
        double vol = F.vol[i];
        double r[5];
        for (int q = 0; q < 5; ++q) r[q] = (double)(R(0.5) * acc[q]);
        r[4] = r[4] - vol * (double)s5;
        if (K.iaxis == 1) r[2] = r[2] - vol * (double)s3;
        double a = alpha * (F.dt[i] / vol);
        const double *Q = F.qn;
        for (int q = 0; q < 5; ++q) cold[q] = Q[q * nt + i];
        double dd[5], vs[4];
        for (int q = 0; q < 5; ++q) dd[q] = ld<SD>(F.dd, (size_t)q * n + i);
        for (int q = 0; q < 4; ++q) vs[q] = ld<SD>(F.vis, (size_t)q * n + i);
        cnew[0] = cold[0] - a * (r[0] - dd[0]);
        cnew[1] = cold[1] - a * (r[1] - vs[0] - dd[1]);
        cnew[2] = cold[2] - a * (r[2] - vs[1] - dd[2]);
        cnew[3] = cold[3] - a * (r[3] - vs[2] - dd[3]);
        cnew[4] = cold[4] - a * kCflTurb * (r[4] - vs[3] - dd[4]);
        cnew[4] = cnew[0] * dmax(K.rtol, cnew[4] / cnew[0]);
        if (STAGE < 4) {
            void *qo = F.st[(STAGE - 1) & 1];
            for (int q = 0; q < 5; ++q) stv<SO>(qo, (size_t)q * n + i, cnew[q]);
        } else {
            double *Qw = F.qn;
            for (int q = 0; q < 5; ++q) Qw[q * nt + i] = cnew[q];
        }
    }
    if (STAGE == 2 && !K.local && blockIdx.x == 0 && threadIdx.x == 0)
        *F.dtslot = (unsigned long long)__double_as_longlong(100000.0);
    if (STAGE == 4) {
        // per-iteration L2 / Linf of C - COLD and the non-finite flag: block
        // partials only; the next pre-pass (or k_finalize) folds them, so no
        // grid-wide fence sits in this kernel
        double sq[5], mx[5];
        int bad = 0;
        for (int q = 0; q < 5; ++q) {
            double d = cnew[q] - cold[q];
            sq[q] = active ? d * d : 0.0;
            mx[q] = active ? fabs(d) : 0.0;
            bad |= active && !isfinite(cnew[q]);
        }
        block_partials_write(sq, mx, bad, F.part + (size_t)blockIdx.x * kPartStride);
        if (blockIdx.x == 0 && threadIdx.x == 0) F.ctr[3] = 1;   // partials pending
    }
}

__global__ void __launch_bounds__(kBlock) k_finalize(FDev F, int nblk) { fold_monitor(F, nblk); }

}  // namespace
}  // namespace tdt

#include "quad_kernels.hpp"

namespace tdt {
namespace {

__global__ void k_reset_counters(int *ctr, unsigned long long *dtslot) {
    ctr[0] = 0;
    ctr[1] = 0;
    ctr[2] = 0;
    ctr[3] = 0;
    *dtslot = (unsigned long long)__double_as_longlong(100000.0);
}

__global__ void __launch_bounds__(kBlock) k_cold_from_qn(FDev F, double *cold) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < F.n)
        for (int q = 0; q < 5; ++q) cold[q * F.n + i] = F.qn[q * F.nt + i];
}

// Launches one kernel; when profiling, with start/stop events that the runtime
// stamps at the kernel's own begin and end (same durations a kernel trace shows).
template <class K, class... Args>
void launch(tdt_ctx *c, FusedBufs &B, int slot, K kern, unsigned grid, Args... args) {
    if (c->profiling && B.have_events)
        hipExtLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, c->stream, B.ev[2 * slot],
                              B.ev[2 * slot + 1], 0, args...);
    else
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, c->stream, args...);
}

template <class R, class SS, class SD>
hipError_t enqueue_iteration(tdt_ctx *c, FusedBufs &B, const Consts<R> &KR) {
    const FDev &F = B.F;
    if (B.quad) {
        const unsigned nq = B.nbq_cells;
        launch(c, B, 0, kq_pre<SD, R>, nq + B.nbq_nodes, F, c->K, KR, B.nbq_cells, B.nbq_cells);
        if (c->profiling && B.have_events) {   // no separate dissipation/viscous kernel
            (void)hipEventRecord(B.ev[2], c->stream);
            (void)hipEventRecord(B.ev[3], c->stream);
        }
        launch(c, B, 2, kq_stage<R, double, SS, SD, 1>, nq, F, c->K, KR, 0.25);
        launch(c, B, 3, kq_stage<R, SS, SS, SD, 2>, nq, F, c->K, KR, 1.0 / 3.0);
        launch(c, B, 4, kq_stage<R, SS, SS, SD, 3>, nq, F, c->K, KR, 0.5);
        launch(c, B, 5, kq_stage<R, SS, double, SD, 4>, nq, F, c->K, KR, 1.0);
        return hipGetLastError();
    }
    const unsigned nbc = B.nb_cells;
    launch(c, B, 0, k_pre<SD, R>, nbc + B.nb_nodes, F, c->K, KR, B.nb_cells);
    launch(c, B, 1, k_dv<SD, R>, nbc, F, c->K, KR);
    launch(c, B, 2, k_stage<R, double, SS, SD, 1>, nbc, F, c->K, KR, 0.25);
    launch(c, B, 3, k_stage<R, SS, SS, SD, 2>, nbc, F, c->K, KR, 1.0 / 3.0);
    launch(c, B, 4, k_stage<R, SS, SS, SD, 3>, nbc, F, c->K, KR, 0.5);
    launch(c, B, 5, k_stage<R, SS, double, SD, 4>, nbc, F, c->K, KR, 1.0);
    return hipGetLastError();
}

hipError_t enqueue(tdt_ctx *c, FusedBufs &B) {
    if (c->precision == TDT_FP64) return enqueue_iteration<double, double, double>(c, B, c->K);
    return enqueue_iteration<float, float, float>(c, B, c->Kf);
}

template <class T>
T *dalloc(FusedBufs &B, size_t count) {
    void *p = nullptr;
    if (hipMalloc(&p, count * sizeof(T) + 256) != hipSuccess) return nullptr;
    B.own.push_back(p);
    return static_cast<T *>(p);
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
    F.gparent = D.gparent; F.gtype = D.gtype; F.gopp = D.gopp;
    F.nptr = D.nptr; F.nlist = D.nlist; F.novr = D.novr;
    F.xcc = D.xcc; F.ycc = D.ycc; F.vol = D.vol; F.dl = D.dl; F.sc = D.sc;
    F.qn = D.c;
    F.dt = D.dt;
    bool ok = true;
    if (c->precision == TDT_FP64) {
        // the routine-level arrays double as the fused path's storage
        F.anu = D.anu; F.anurt = D.anurt; F.d2q = D.d2q; F.dd = D.d24; F.vis = D.vis;
        F.node = D.nodeavg;
        F.st[0] = dalloc<double>(*B, (size_t)5 * D.n);
        F.st[1] = dalloc<double>(*B, (size_t)5 * D.n);
        ok = F.st[0] && F.st[1];
    } else {
        F.anu = dalloc<float>(*B, D.n);
        F.anurt = dalloc<float>(*B, D.n);
        F.d2q = dalloc<float>(*B, (size_t)5 * D.nt);
        F.dd = dalloc<float>(*B, (size_t)5 * D.n);
        F.vis = dalloc<float>(*B, (size_t)4 * D.n);
        F.node = dalloc<float>(*B, (size_t)5 * D.m);
        F.st[0] = dalloc<float>(*B, (size_t)5 * D.n);
        F.st[1] = dalloc<float>(*B, (size_t)5 * D.n);
        ok = F.anu && F.anurt && F.d2q && F.dd && F.vis && F.node && F.st[0] && F.st[1];
    }
    B->nb_cells = (int)nblocks(D.n);
    B->nb_nodes = (int)nblocks(D.m);
    B->nbq_cells = (D.n + kBlock / 4 - 1) / (kBlock / 4);
    B->nbq_nodes = (D.m + kBlock / 4 - 1) / (kBlock / 4);
    const char *kv = std::getenv("TDT_KERNELS");
    B->quad = kv && std::string(kv) == "quad";
    F.dtslot = dalloc<unsigned long long>(*B, 1);
    F.ctr = dalloc<int>(*B, 4);
    F.part = dalloc<double>(*B, (size_t)std::max(B->nb_cells, B->nbq_cells) * kPartStride);
    ok = ok && F.dtslot && F.ctr && F.part;
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

int fused_iterate(tdt_ctx *c, int niter, double *l2, double *linf, int *first_nonfinite) {
    int rc = fused_setup(c);
    if (rc != TDT_OK) return rc;
    FusedBufs &B = *c->fused;
    FDev &F = B.F;
    if ((size_t)niter > B.hist_cap) {
        size_t cap = std::max((size_t)niter, (size_t)1024);
        double *h = dalloc<double>(B, cap * 10);
        if (!h) {
            c->err = "device allocation of the monitor history failed";
            return TDT_ERR_DEVICE;
        }
        F.hist = h;
        B.hist_cap = cap;
        // kernels captured earlier point at the old history buffer
        if (B.exec1) { hipGraphExecDestroy(B.exec1); B.exec1 = nullptr; }
        if (B.execK) { hipGraphExecDestroy(B.execK); B.execK = nullptr; }
        if (B.graph1) { hipGraphDestroy(B.graph1); B.graph1 = nullptr; }
        if (B.graphK) { hipGraphDestroy(B.graphK); B.graphK = nullptr; }
    }
    hipError_t e = hipSuccess;
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
    } else if (e == hipSuccess) {
        if (!B.exec1) e = build_graph(c, B, 1, &B.graph1, &B.exec1);
        if (e == hipSuccess && !B.execK && niter >= B.chunk)
            e = build_graph(c, B, B.chunk, &B.graphK, &B.execK);
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
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, c->stream, F,
                           B.quad ? B.nbq_cells : B.nb_cells);
        hipLaunchKernelGGL(k_cold_from_qn, dim3(nblocks(F.n)), dim3(kBlock), 0, c->stream, F,
                           c->D.cold);
    }
    if (e == hipSuccess) e = hipGetLastError();
    int ctr[4] = {0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(ctr, F.ctr, sizeof ctr, hipMemcpyDeviceToHost, c->stream);
    std::vector<double> h;
    if (e == hipSuccess && (l2 || linf) && niter > 0) {
        h.resize((size_t)niter * 10);
        e = hipMemcpyAsync(h.data(), F.hist, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        c->err = std::string("fused iteration: ") + hipGetErrorString(e);
        return TDT_ERR_DEVICE;
    }
    for (int it = 0; it < (int)h.size() / 10; ++it)
        for (int q = 0; q < 5; ++q) {
            if (l2) l2[(size_t)it * 5 + q] = h[(size_t)it * 10 + q];
            if (linf) linf[(size_t)it * 5 + q] = h[(size_t)it * 10 + 5 + q];
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
