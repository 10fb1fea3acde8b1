// Face-parallel fused kernels: a quad of four lanes per cell, lane f owning
// face f (its neighbour's primitives or ghost, flux terms, dissipation face
// terms, viscous face).  Sums over faces are formed with DPP quad moves in the
// reference's face order, so every result is bit-identical to the one-lane-
// per-cell kernels; what changes is that four times as many wavefronts are in
// flight and the per-face transcendentals (viscous pow / tanh, the four
// source-term tanh, the gradient divisions) run side by side instead of one
// after another.  Included by fused.hip (shares FDev, ld / stv, fold_monitor).
#pragma once

#include "quad.hpp"

namespace tdt {
namespace {

constexpr int kCellsPerBlock = kBlock / 4;

// ---------------------------------------------------------------------------------
// kq_pre: blocks [0, nbq_cells) do cells, the rest do nodes (64 per block); the
// last block folds the previous iteration's monitor partials.
template <class SD, class R, bool B6>
__global__ void __launch_bounds__(kBlock) kq_pre(FDev F, Consts<double> K, Consts<R> KR,
                                                 int nbq_cells, int nb_partials) {
    const int lane = threadIdx.x & 3;
    const int b = blockIdx.x;
    if (b == (int)gridDim.x - 1) {
        fold_monitor(F, nb_partials);
        return;
    }
    const int nt = F.nt, n = F.n;
    const double *Q = F.qn;
    if (b < nbq_cells) {
        const int cq = b * kCellsPerBlock + (threadIdx.x >> 2);
        const bool act = cq < n;
        const int i = act ? cq : n - 1;
        double c0 = Q[i], c1 = Q[nt + i], c2 = Q[2 * nt + i], c3 = Q[3 * nt + i],
               c4 = Q[4 * nt + i];
        double u, v, t, p;
        uvtp_from_q(c0, c1, c2, c3, u, v, t, p, K);
        double vel = sqrt(u * u + v * v);
        double dt = fdiv(K.cfl * F.dl[i], fabs(vel) + sqrt(K.gamma_rr * t));
        if (act && lane == 0 && K.local) F.dt[i] = dt;
        // the neighbour across this lane's face; a ghost holds the parent's copy
        // in DISSIPATION (solver_state.f90:77-88), i.e. this cell
        const int j = reinterpret_cast<const int *>(F.nc4)[4 * i + lane];
        const int jj = j < n ? j : i;
        double q0 = Q[jj], q1 = Q[nt + jj], q2 = Q[2 * nt + jj], q3 = Q[3 * nt + jj],
               q4 = Q[4 * nt + jj];
        double uj, vj, tj, pj;
        uvtp_from_q(q0, q1, q2, q3, uj, vj, tj, pj, K);
        double anun, anud, rtn, rtd;
        qabs_sums(pj, p, anun, anud);
        qabs_sums(q4, c4, rtn, rtd);
        double d2q[5];
        d2q[0] = qsum(q0) - 4.0 * c0;
        d2q[1] = qsum(q1) - 4.0 * c1;
        d2q[2] = qsum(q2) - 4.0 * c2;
        d2q[3] = qsum(q3) - 4.0 * c3;
        d2q[4] = qsum(q4) - 4.0 * c4;
        if (act) {
            // this cell's record: lane e writes D2Q(e+1); lane 0 also D2Q5 and
            // ANU, lane 1 ANURT, lane 2 VOL/DT (VOL when the step is global)
            rec_slot<SD>(F.drec, n, i, kRecD2q + lane) = (SD)sel4(lane, d2q[0], d2q[1], d2q[2], d2q[3]);
            if (lane == 0) {
                rec_slot<SD>(F.drec, n, i, kRecD2q + 4) = (SD)d2q[4];
                rec_slot<SD>(F.drec, n, i, kRecAnu) = (SD)fdiv(anun, anud);
            } else if (lane == 1) {
                rec_slot<SD>(F.drec, n, i, kRecAnurt) = (SD)fdiv(rtn, dmax(rtd, 8.0 * c0 * K.rtol));
            } else if (lane == 2) {
                const R vol = (R)F.vol[i];
                rec_slot<SD>(F.drec, n, i, kRecVbydt) = (SD)(K.local ? fdiv(vol, (R)dt) : vol);
            }
        }
        if (!K.local) block_atomic_min_pos(act ? dt : 100000.0, reinterpret_cast<double *>(F.dtslot));
        return;
    }
    // nodal averages (viscous.f90:40-124): the quad walks the node's contributors four
    // at a time and accumulates them in list order
    const int vq = (b - nbq_cells) * kCellsPerBlock + (threadIdx.x >> 2);
    const bool actv = vq < F.m;
    const int v = actv ? vq : F.m - 1;
    const int4 ov = F.novr[v];
    const int j0 = F.nptr[v], j1 = F.nptr[v + 1];
    R au = R(0), av = R(0), at = R(0), arho = R(0), ar = ov.w ? KR.rtol_rtred : R(0);
    for (int base = j0; base < j1; base += 4) {
        const int jl = base + lane;
        const int id = F.nlist[jl < j1 ? jl : base];
        R uu, vv, tt, rho, c5;
        if (id < n) {
            R pp;
            rho = (R)Q[id];
            uvtp_from_q<R>(rho, (R)Q[nt + id], (R)Q[2 * nt + id], (R)Q[3 * nt + id], uu, vv, tt, pp, KR);
            c5 = (R)Q[4 * nt + id];
        } else {
            const int k = id - n, par = F.gparent[k], ty = F.gtype[k];
            Prim<R> pp, s;
            prim_from_q<R>(pp, (R)Q[par], (R)Q[nt + par], (R)Q[2 * nt + par], (R)Q[3 * nt + par],
                           (R)Q[4 * nt + par], KR);
            R uo = R(0), vo = R(0), to = R(0), po = R(0);
            if (ty == 3) {
                const int o = F.gopp[k];
                uvtp_from_q<R>((R)Q[o], (R)Q[nt + o], (R)Q[2 * nt + o], (R)Q[3 * nt + o], uo, vo,
                               to, po, KR);
            }
            if (B6 && ty == 6) {
                bc6_load<R>(F, k, uu, vv, tt, rho, c5);
            } else {
                ghost_state<R>(s, ty, pp, uo, vo, to, po, (R)F.ycc[id], KR);
                uu = s.u; vv = s.v; tt = s.t; rho = s.c[0]; c5 = s.c[4];
            }
        }
        const R rr = fdiv(c5, rho);
#define TDT_NODE_ACC(K_)                                                           \
        {                                                                          \
            const int pos = base + K_;                                             \
            const R xu = qget<K_>(uu), xv = qget<K_>(vv), xt = qget<K_>(tt);       \
            const R xrho = qget<K_>(rho), xr = qget<K_>(rr);                       \
            if (pos < j1) {                                                        \
                if (pos >= ov.x) au += xu;                                         \
                if (pos >= ov.y) av += xv;                                         \
                at += xt;                                                          \
                arho += xrho;                                                      \
                if (pos >= ov.z) ar += xr;                                         \
            }                                                                      \
        }
        TDT_NODE_ACC(0) TDT_NODE_ACC(1) TDT_NODE_ACC(2) TDT_NODE_ACC(3)
#undef TDT_NODE_ACC
    }
    if (actv) {
        // this node's averages: lane e writes average e, lane 0 also RN
        const R cnt = (R)(j1 - j0);
        const Recip<R> rcnt(cnt);
        const R val = rcnt.div(sel4(lane, au, av, at, arho));
        const int m = F.m;
        nav_slot<SD>(F.nrec, m, v, lane) = (SD)val;
        if (lane == 0) nav_slot<SD>(F.nrec, m, v, 4) = (SD)dmax(KR.rtol, rcnt.div(ar));
    }
}

// ---------------------------------------------------------------------------------
// kq_stage: one RK stage; stage 1 also forms the frozen dissipation D and
// viscous fluxes VIS (solver_state.f90:120-223, viscous.f90:62-217).
template <class R, class SI, class SO, class SD, int STAGE, bool B6>
__global__ void __launch_bounds__(kBlock) kq_stage(FDev F, Consts<double> K, Consts<R> KR,
                                                   double alpha) {
    const int lane = threadIdx.x & 3;
    const int cq = logical_block(F.remap) * kCellsPerBlock + (threadIdx.x >> 2);
    const int nt = F.nt, n = F.n;
    const bool act = cq < n;
    const int i = act ? cq : n - 1;
    const void *qin = STAGE == 1 ? (const void *)F.qn : (const void *)F.st[STAGE & 1];
    const size_t sin = STAGE == 1 ? (size_t)nt : (size_t)n;
    const double *Q = F.qn;

    Prim<R> pi;
    prim_from_q<R>(pi, (R)ld<SI>(qin, i), (R)ld<SI>(qin, sin + i), (R)ld<SI>(qin, 2 * sin + i),
                   (R)ld<SI>(qin, 3 * sin + i), (R)ld<SI>(qin, 4 * sin + i), KR);
    const int j = reinterpret_cast<const int *>(F.nc4)[4 * i + lane];
    const bool ghost = j >= n;
    const int jj = ghost ? i : j;
    Prim<R> pj;
    prim_from_q<R>(pj, (R)ld<SI>(qin, jj), (R)ld<SI>(qin, sin + jj), (R)ld<SI>(qin, 2 * sin + jj),
                   (R)ld<SI>(qin, 3 * sin + jj), (R)ld<SI>(qin, 4 * sin + jj), KR);
    // the parent's opposite neighbour, needed by type-3 ghosts, is the neighbour
    // the lane across the cell just loaded
    const R uo = qopp(pj.u), vo = qopp(pj.v), to = qopp(pj.t), po = qopp(pj.p);
    if (ghost)
        ghost_state<R>(pj, F.gtype[j - n], pi, uo, vo, to, po, (R)F.ycc[j], KR,
                       g6_for<B6, R>(F, j - n));

    // central flux residual (residual.f90:34-54): face terms summed in face order
    const R sx = (R)F.sc[(2 * lane) * n + i], sy = (R)F.sc[(2 * lane + 1) * n + i];
    R r[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) r[q] = R(0.5) * qsum2((pi.f[q] + pj.f[q]) * sx, (pi.g[q] + pj.g[q]) * sy);

    const R xj = (R)F.xcc[j], yj = (R)F.ycc[j];
    double dcell[5], vcell[4];
    if (STAGE == 1) {
        // dissipation face terms; a ghost neighbour is a copy of this cell
        const double dti = K.local ? F.dt[i] : __longlong_as_double((long long)*F.dtslot);
        SD ri[kRecLen], rj[kRecLen];
        rec_get<SD>(F.drec, n, i, ri);
        rec_get<SD>(F.drec, n, jj, rj);
        const R volbydt = rec_vbydt<R>(ri, K.local, dti);
        const R vbn = rec_vbydt<R>(rj, K.local, dti);
        R a2, a4, b2, b4;
        jst_coeffs_v<R>((R)ri[kRecAnu], (R)rj[kRecAnu], volbydt, vbn, a2, a4);
        jst_coeffs_v<R>((R)ri[kRecAnurt], (R)rj[kRecAnurt], volbydt, vbn, b2, b4);
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const R c2 = q < 4 ? a2 : b2, c4 = q < 4 ? a4 : b4;
            const R dc = (R)(Q[q * nt + jj] - Q[q * nt + i]);
            const R dq = (R)rj[kRecD2q + q] - (R)ri[kRecD2q + q];
            dcell[q] = (double)(SD)(qsum(c2 * dc) - qsum(c4 * dq));   // as stored for stages 2-4
        }
        // viscous face (viscous.f90:130-217) with nodal averages of this face's nodes
        const int4 nd4 = F.nod4[i];
        const int bnode = sel4(lane, nd4.x, nd4.y, nd4.z, nd4.w);
        const int dnode = sel4(lane, nd4.y, nd4.z, nd4.w, nd4.x);
        const R ra = fdiv(pi.c[4], pi.c[0]), rc = fdiv(pj.c[4], pj.c[0]);
        ViscFace<R> o = viscous_face_rec<R, SD>(F, sx, sy, (R)F.xcc[i], (R)F.ycc[i], pi.u, pi.v,
                                                pi.t, pi.c[0], ra, bnode, dnode, xj, yj, pj.u,
                                                pj.v, pj.t, pj.c[0], rc, KR);
        vcell[0] = (double)(SD)qsum2(o.u0, o.u1);
        vcell[1] = (double)(SD)qsum2(o.v0, o.v1);
        vcell[2] = (double)(SD)qsum2(o.t0, o.t1);
        vcell[3] = (double)(SD)qsum2(o.r0, o.r1);
    }

    // source terms (source_terms.f90:40-101): gradients of |V|, nu~, u, v from the
    // four neighbour centroids; lane f forms the x/y gradient of quantity f
    const R xa = qget<0>(xj), xb = qget<1>(xj), xc = qget<2>(xj), xd = qget<3>(xj);
    const R ya = qget<0>(yj), yb = qget<1>(yj), yc = qget<2>(yj), yd = qget<3>(yj);
    const R rj = fdiv(pj.c[4], pj.c[0]);
    const R qj = tsqrt(pj.u * pj.u + pj.v * pj.v);
    const R wa = sel4(lane, qget<0>(qj), qget<0>(rj), qget<0>(pj.u), qget<0>(pj.v));
    const R wb = sel4(lane, qget<1>(qj), qget<1>(rj), qget<1>(pj.u), qget<1>(pj.v));
    const R wc = sel4(lane, qget<2>(qj), qget<2>(rj), qget<2>(pj.u), qget<2>(pj.v));
    const R wd = sel4(lane, qget<3>(qj), qget<3>(rj), qget<3>(pj.u), qget<3>(pj.v));
    const R dnr = (xa - xc) * (yb - yd) - (xb - xd) * (ya - yc);
    const Recip<R> rdnr(dnr);
    const R gx = rdnr.div((wa - wc) * (yb - yd) - (wb - wd) * (ya - yc));
    const R gy = rdnr.div(-((wa - wc) * (xb - xd)) + (wb - wd) * (xa - xc));
    const R dqx = qget<0>(gx), dqy = qget<0>(gy), drx = qget<1>(gx), dry = qget<1>(gy);
    const R dux = qget<2>(gx), duy = qget<2>(gy), dvx = qget<3>(gx), dvy = qget<3>(gy);
    const R rho = pi.c[0];
    const Recip<R> rrho(rho);
    const R ri = rrho.div(pi.c[4]);
    const R amu = KR.suth_sqrt_amol * tpow(pi.t, R(0.6));
    const R chi = fdiv(rho * ri, R(kCmu) * amu);
    // the four tanh of FMU and FR2, one per lane
    const R targ = sel4(lane, R(kAlphaTurb) * chi * chi, R(kBetaTurb) * chi * chi, tsqrt(chi),
                        chi * KR.sqcmu);
    const R th = ttanh(targ);
    const R fmu = fdiv(qget<0>(th), qget<1>(th));
    const R fr2 = fdiv(qget<2>(th), qget<3>(th));
    const R amut = fmu * rho * ri;
    const R phi = dqx * drx + dqy * dry;
    R dee = R(0);
    if (phi > R(0)) dee = drx * drx + dry * dry;
    const R ratio = ipow(fdiv(amut, dmax(amut, R(kAmuThresh) * amu)), KR.npseudo);
    R s5, s3 = R(0);
    if (KR.iaxis == 1) {
        const R ycci = (R)F.ycc[i];
        const Recip<R> rycc(ycci);
        const R vbyy = (ycci < R(kYaxisThresh)) ? dvy : rycc.div(pi.v);
        const R divvel = dux + dvy + vbyy;
        R pk = rrho.div(amut) *
               (dux * dux + dvy * dvy + vbyy * vbyy + dux * dux + dvy * dvy + vbyy * vbyy +
                (duy + dvx) * (duy + dvx) - R(kDiv23) * divvel * divvel);
        pk = dmax(pk, R(0));
        s5 = (rho * (R(kConst1) - R(kConst2) * fr2) * tsqrt(ri * pk) - rho * R(kConst3) * dee) * ratio;
        const R ttt = (amu + amut) * (vbyy + vbyy - R(kDiv23) * divvel);
        s3 = rycc.div(pi.p - ttt);
    } else {
        const R divvel = dux + dvy;
        R pk = rrho.div(amut) *
               (dux * dux + dvy * dvy + dux * dux + dvy * dvy + (duy + dvx) * (duy + dvx) -
                R(kDiv23) * divvel * divvel);
        pk = dmax(pk, R(0));
        s5 = (rho * (R(kConst1) - R(kConst2) * fr2) * tsqrt(ri * pk) - rho * R(kConst3) * dee) * ratio;
    }

    // update (residual.f90:67-83) in FP64: lane e advances equation e, lane 0 also
    // equation 5 (it needs the new rho for the nu~ clamp)
    const double vol = F.vol[i];
    double rr[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) rr[q] = (double)r[q];
    rr[4] = rr[4] - vol * (double)s5;
    if (K.iaxis == 1) rr[2] = rr[2] - vol * (double)s3;
    double dti;
    if (STAGE == 1) {
        dti = K.local ? F.dt[i] : __longlong_as_double((long long)*F.dtslot);
        if (act && lane == 0 && !K.local) F.dt[i] = dti;
    } else {
        dti = F.dt[i];
    }
    const double a = alpha * fdiv(dti, vol);
    const int e = lane;
    const double coldE = Q[e * nt + i];
    double dE, visE;
    constexpr bool kE = kCombinedDV<SD>;   // FP32-mixed: E = D + VIS stored in D
    if (STAGE == 1) {
        dE = sel4(e, dcell[0], dcell[1], dcell[2], dcell[3]);
        visE = sel4(e, 0.0, vcell[0], vcell[1], vcell[2]);
        if (kE) {
            dE = (double)(SD)(dE + visE);
            visE = 0.0;
        }
    } else {
        dE = ld<SD>(F.dd, (size_t)e * n + i);
        visE = e == 0 || kE ? 0.0 : ld<SD>(F.vis, (size_t)(e - 1) * n + i);
    }
    const double rE = sel4(e, rr[0], rr[1], rr[2], rr[3]);
    double cE = e == 0 ? coldE - a * (rE - dE) : coldE - a * (rE - visE - dE);
    double cold4 = 0.0, c4 = 0.0, e4 = 0.0;
    if (lane == 0) {
        cold4 = Q[4 * nt + i];
        double d4 = STAGE == 1 ? dcell[4] : ld<SD>(F.dd, (size_t)4 * n + i);
        double v4 = STAGE == 1 ? vcell[3] : (kE ? 0.0 : ld<SD>(F.vis, (size_t)3 * n + i));
        if (kE && STAGE == 1) {
            d4 = (double)(SD)(d4 + v4);
            v4 = 0.0;
        }
        e4 = d4;
        c4 = cold4 - a * kCflTurb * (rr[4] - v4 - d4);
        c4 = cE * dmax(K.rtol, fdiv(c4, cE));
    }
    if (act) {
        if (STAGE == 1) {
            stv<SD>(F.dd, (size_t)e * n + i, dE);
            if (!kE) stv<SD>(F.vis, (size_t)e * n + i, sel4(e, vcell[0], vcell[1], vcell[2], vcell[3]));
            if (lane == 0) stv<SD>(F.dd, (size_t)4 * n + i, e4);
        }
        if (STAGE < 4) {
            void *qo = F.st[(STAGE - 1) & 1];
            stv<SO>(qo, (size_t)e * n + i, cE);
            if (lane == 0) stv<SO>(qo, (size_t)4 * n + i, c4);
        } else {
            F.qn[e * nt + i] = cE;
            if (lane == 0) F.qn[4 * nt + i] = c4;
        }
    }
    if (STAGE == 2 && !K.local && blockIdx.x == 0 && threadIdx.x == 0)
        *F.dtslot = (unsigned long long)__double_as_longlong(100000.0);
    if (STAGE == 4) {
        double sq[5] = {0, 0, 0, 0, 0}, mx[5] = {0, 0, 0, 0, 0};
        int bad = 0;
        if (act) {
            const double d = cE - coldE;
            const double d4 = c4 - cold4;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                sq[q] = q == e ? d * d : 0.0;
                mx[q] = q == e ? fabs(d) : 0.0;
            }
            sq[4] = lane == 0 ? d4 * d4 : 0.0;
            mx[4] = lane == 0 ? fabs(d4) : 0.0;
            bad = !isfinite(cE) || (lane == 0 && !isfinite(c4));
        }
        block_partials_write(sq, mx, bad, F.part + (size_t)blockIdx.x * kPartStride);
        if (blockIdx.x == 0 && threadIdx.x == 0) F.ctr[3] = 1;
    }
}

}  // namespace
}  // namespace tdt
