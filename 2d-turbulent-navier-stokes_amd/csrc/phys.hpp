// Device-side physics of the residual + RK4 path.  Each function restates one
// piece of the reference solver with the reference's evaluation order
// (Fortran: left to right within a precedence level, parentheses honoured).
// With R = double and -ffp-contract=off the arithmetic is bit-identical to the
// reference except where libm (pow, tanh) differs from the device math
// library.  R = float is the FP32-mixed arithmetic.
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>

#include "tmath.hpp"

namespace tdt {

// Model and scheme constants (constants.f90:15-45).
constexpr double kCmu = 0.09, kAlphaTurb = 0.00052, kBetaTurb = 0.2;
constexpr double kConst1 = 21.479, kConst2 = 21.394, kConst3 = 1.5;
constexpr double kSuth = 0.00000011848;
constexpr double kC02 = 0.25, kC04 = 0.00390625;
constexpr double kCflTurb = 0.5;
constexpr double kDiv23 = 2.0 / 3.0;
constexpr double kYaxisThresh = 0.001, kAmuThresh = 0.001;
constexpr double kBlExponent = 1.0 / 7.0;
constexpr double kPaToBar = 0.00001;
constexpr double kConvKsc = 1.0 / 98066.5;   // Pa -> kgf/cm^2 (constants.f90:55)
constexpr double kGth = 35.0;                // erosive-burning threshold (constants.f90:56)

// Flow constants derived once on the host with the reference's expressions.
template <class R>
struct Consts {
    R gama, cfl, prandtl, prandtlt, t0, pamb, pinf, deltabl;
    R rr, cv, cp, s0, tinf;
    R rtol, rtol_rtred;        // RTOLERANCE and RTOLERANCE*RTRED
    R suth_sqrt_amol;          // SUTH_COEF*DSQRT(AMOL)
    R inv_1mg;                 // 1/(1-GAMA)
    R sqcmu;                   // DSQRT(2*DSQRT(CMU))
    R bl_u;                    // 2*AMINF*DSQRT(GAMA*RR*TINF)
    R gamma_rr;                // GAMA*RR
    R tflame, ancomb, acomb, rhoprop;  // BC type 6 inputs
    R gth_08;                  // GTH**0.8 (constants.f90:56), from the host's libm
    int local, npseudo, iaxis, pamb_set;
};

template <class R>
__device__ __forceinline__ R dmax(R a, R b) { return a > b ? a : b; }
template <class R>
__device__ __forceinline__ R dmin(R a, R b) { return a < b ? a : b; }

// FP64: the branch-free functions of tmath.hpp (a few ulps of glibc).  FP32
// (mixed path): pow as 2^(b log2 a) with the hardware log2 / exp2 (a few ulps),
// tanh from the device library.
__device__ __forceinline__ double tpow(double a, double b) { return tm::pow(a, b); }
__device__ __forceinline__ float tpow(float a, float b) {
    return __builtin_amdgcn_exp2f(b * __builtin_amdgcn_logf(a));
}
__device__ __forceinline__ double ttanh(double a) { return tm::tanh(a); }
__device__ __forceinline__ float ttanh(float a) { return tanhf(a); }
__device__ __forceinline__ double tsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float tsqrt(float a) { return sqrtf(a); }

// Division by a reused denominator.  FP64 division on gfx950 is r = rcp(b),
// two Newton steps r += r (1 - b r), q = a r, q += r (a - b q) (Markstein's
// correction, correctly rounded given the refined r), wrapped in operand
// pre-scaling (v_div_scale), post-scaling (v_div_fmas) and a special-case
// fix-up (v_div_fixup) that change nothing unless an operand is zero, infinite,
// NaN or denormal, the quotient is denormal or overflows, or the exponents are
// about 768 apart.  Recip keeps the refined reciprocal of b, so each numerator
// costs a multiply and two FMAs instead of the whole sequence; for the finite,
// normal operands of a physical state the result has the same bits as a / b.
// Outside that range: a / 0, inf / b and b = inf give NaN (a / b: inf or 0),
// still non-finite where a / b is; -0 / b gives +0 for b > 0.
// FP32 (mixed path): plain a / b, compiled there with reciprocal math.
template <class R>
struct Recip;
template <>
struct Recip<double> {
    double b, r;
    __device__ __forceinline__ explicit Recip(double den) : b(den) {
        const double r0 = __builtin_amdgcn_rcp(den);
        const double r1 = __builtin_fma(r0, __builtin_fma(-den, r0, 1.0), r0);
        r = __builtin_fma(r1, __builtin_fma(-den, r1, 1.0), r1);
    }
    __device__ __forceinline__ double div(double a) const {
        const double q = a * r;
        return __builtin_fma(__builtin_fma(-b, q, a), r, q);
    }
};
template <>
struct Recip<float> {
    float b;
    __device__ __forceinline__ explicit Recip(float den) : b(den) {}
    __device__ __forceinline__ float div(float a) const { return a / b; }
};

// a / b through Recip: same bits as a / b under the conditions above, three
// VALU instructions shorter (FP64)
template <class R>
__device__ __forceinline__ R fdiv(R a, R b) { return Recip<R>(b).div(a); }

// real**integer by binary exponentiation (Fortran's real**integer).
template <class R>
__device__ __forceinline__ R ipow(R a, int b) {
    bool neg = b < 0;
    unsigned ub = neg ? 0u - (unsigned)b : (unsigned)b;
    R r = R(1);
    while (ub) {
        if (ub & 1u) r *= a;
        ub >>= 1;
        if (ub) a *= a;
    }
    return neg ? R(1) / r : r;
}

// Cell state with primitives and fluxes (COMPUTE_PRIMITIVE_VARS outputs).
template <class R>
struct Prim {
    R c[5];
    R u, v, t, p;
    R f[5], g[5];
};

// turbnstwod.f90:204-227
template <class R>
__device__ __forceinline__ void prim_from_q(Prim<R> &s, R q0, R q1, R q2, R q3, R q4,
                                            const Consts<R> &K) {
    R rho = q0;
    s.c[0] = q0; s.c[1] = q1; s.c[2] = q2; s.c[3] = q3; s.c[4] = q4;
    const Recip<R> rr(rho);
    s.u = rr.div(q1);
    s.v = rr.div(q2);
    R e = rr.div(q3);
    s.t = Recip<R>(K.cv).div(e - R(0.5) * (s.u * s.u + s.v * s.v));
    s.p = rho * K.rr * s.t;
    s.f[0] = q1;
    s.f[1] = rho * s.u * s.u + s.p;
    s.f[2] = rho * s.u * s.v;
    s.f[3] = (q3 + s.p) * s.u;
    s.g[0] = q2;
    s.g[1] = s.f[2];
    s.g[2] = rho * s.v * s.v + s.p;
    s.g[3] = (q3 + s.p) * s.v;
    R rt = rr.div(q4);
    s.f[4] = rho * rt * s.u;
    s.g[4] = rho * rt * s.v;
}

// u, v, T, p only (the first lines of turbnstwod.f90:97-114 / 204-227).
template <class R>
__device__ __forceinline__ void uvtp_from_q(R q0, R q1, R q2, R q3, R &u, R &v, R &t, R &p,
                                            const Consts<R> &K) {
    R rho = q0;
    const Recip<R> rr(rho);
    u = rr.div(q1);
    v = rr.div(q2);
    R e = rr.div(q3);
    t = Recip<R>(K.cv).div(e - R(0.5) * (u * u + v * v));
    p = rho * K.rr * t;
}

// Common tail of boundary.f90:62-69, 89-93, 103-107, 321-325.
template <class R>
__device__ __forceinline__ void ghost_fill(Prim<R> &s, R rho, const Consts<R> &K) {
    s.c[0] = rho;
    s.c[1] = rho * s.u;
    s.c[2] = rho * s.v;
    R e = K.cv * s.t + R(0.5) * (s.u * s.u + s.v * s.v);
    s.c[3] = rho * e;
    s.f[0] = s.c[1];
    s.f[1] = rho * s.u * s.u + s.p;
    s.f[2] = rho * s.u * s.v;
    s.f[3] = (s.c[3] + s.p) * s.u;
    s.g[0] = s.c[2];
    s.g[1] = s.f[2];
    s.g[2] = rho * s.v * s.v + s.p;
    s.g[3] = (s.c[3] + s.p) * s.v;
}

// Common tail of the ghosts that get a new state (types 2, 6, 7): C, F, G from
// the primitives in s (u, v, t, p), the density and C5 (boundary.f90:62-70,
// 157-162, 321-326).
template <class R>
__device__ __forceinline__ void ghost_complete(Prim<R> &s, R rho, R c5, const Consts<R> &K) {
    ghost_fill(s, rho, K);
    s.c[4] = c5;
    s.f[4] = s.c[4] * s.u;
    s.g[4] = s.c[4] * s.v;
}

// Type-6 ghosts come from elsewhere (a port-column walk): callers that meet
// them pass a loader g6(u, v, t, p, rho, c5); with this one (grids without
// type 6, or callers that handle it themselves) the last branch is type 7.
struct NoBc6 {
    template <class R>
    __device__ void operator()(R &u, R &v, R &t, R &p, R &rho, R &c5) const {
        u = v = t = p = rho = c5 = R(0);
    }
};

// Ghost-cell state for boundary types 2, 3, 4, 5, 7 (boundary.f90:53-122, 310-327),
// and 6 through g6.  par: parent state; (uo, vo, to, po): primitives of the
// parent's opposite neighbour (type 3 only); ycc_g: ghost centroid y (type 7 only).
template <class R, class G6 = NoBc6>
__device__ __forceinline__ void ghost_state(Prim<R> &s, int type, const Prim<R> &par, R uo, R vo,
                                            R to, R po, R ycc_g, const Consts<R> &K,
                                            G6 g6 = G6{}) {
    if (type == 2) {                       // subsonic isentropic inlet
        s.u = fabs(par.u);
        s.v = -par.v;
        s.t = K.t0 - R(0.5) * (s.u * s.u + s.v * s.v) / K.cp;
        R rho = tpow(K.s0 / (K.rr * s.t), K.inv_1mg);
        s.p = rho * K.rr * s.t;
        R amu = K.suth_sqrt_amol * tpow(s.t, R(0.6));
        ghost_complete(s, rho, rho * (amu / rho), K);
    } else if (type == 3) {                // extrapolation "farfield"
        s.u = fabs(R(2) * par.u - uo);
        s.v = R(2) * par.v - vo;
        s.t = R(2) * par.t - to;
        s.p = R(2) * par.p - po;
        if (K.pamb_set) s.p = K.pamb;
        R rho = s.p / (K.rr * s.t);
        ghost_fill(s, rho, K);
        s.c[4] = par.c[4];
        s.f[4] = par.f[4];
        s.g[4] = par.g[4];
    } else if (type == 4) {                // symmetry
        s.u = par.u;
        s.v = -par.v;
        s.t = par.t;
        s.p = par.p;
        ghost_fill(s, par.c[0], K);
        s.c[4] = par.c[4];
        s.f[4] = par.f[4];
        s.g[4] = -par.g[4];
    } else if (type == 5) {                // no-slip wall / axis
        s.u = -par.u;
        s.v = -par.v;
        s.t = par.t;
        s.p = par.p;
        R rho = par.c[0];
        s.c[0] = rho;
        s.c[1] = rho * s.u;
        s.c[2] = rho * s.v;
        R e = K.cv * s.t + R(0.5) * (s.u * s.u + s.v * s.v);
        s.c[3] = rho * e;
        s.f[0] = -par.f[0];
        s.f[1] = R(2) * par.p - par.f[1];
        s.f[2] = -par.f[2];
        s.f[3] = -par.f[3];
        s.g[0] = -par.g[0];
        s.g[1] = -par.g[1];
        s.g[2] = R(2) * par.p - par.g[2];
        s.g[3] = -par.g[3];
        s.c[4] = -par.c[4];
        s.f[4] = -par.f[4];
        s.g[4] = -par.g[4];
    } else {                               // type 7, and 6 when g6 is a loader
        R rho, c5;
        if (std::is_same<G6, NoBc6>::value || type == 7) {   // 1/7-power boundary-layer inlet
            s.p = K.pinf;
            s.t = K.tinf;
            s.v = -par.v;
            rho = s.p / (K.rr * s.t);
            if (ycc_g <= K.deltabl)
                s.u = K.bl_u * tpow(ycc_g / K.deltabl, R(kBlExponent)) - par.u;
            else
                s.u = K.bl_u - par.u;
            R amu = K.suth_sqrt_amol * tpow(s.t, R(0.6));
            c5 = rho * (amu / rho);
        } else {                           // 6: burning surface, formed beforehand
            g6(s.u, s.v, s.t, s.p, rho, c5);
        }
        ghost_complete(s, rho, c5, K);
    }
}

// BC type 6 (boundary.f90:128-305): the ghost of a burning propellant surface.
// The port mass flow GPORT sums C1*U*DELY (*YCC when IAXIS = 1) over the column
// of cells from the parent across to the far boundary, entries [e0, e1) of
// cells / dely in the reference's order; cell(n, c1, c2, ycc) loads a column
// cell's C1, C2 and centroid y.  (scx, scy): the parent's surface vector on the
// ghost's face; dhyd = |YCC(ghost) + YCC(parent)|.  The burn rate follows
// ACOMB (P CONVKSC)^ANCOMB with the erosive correction above GVAL = GTH.
template <class R, class Cell>
__device__ __forceinline__ void bc6_state(Prim<R> &s, const Prim<R> &par, R scx, R scy, R dhyd,
                                          int e0, int e1, const int *cells, const double *dely,
                                          Cell cell, const Consts<R> &K) {
    R gp = R(0);
    for (int e = e0; e < e1; ++e) {
        R c1, c2, yc;
        cell(cells[e], c1, c2, yc);
        const R u = Recip<R>(c1).div(c2);          // U = C2/C1 (COMPUTE_PRIMITIVE_VARS)
        const R t = K.iaxis == 1 ? c1 * u * (R)dely[e] * yc : c1 * u * (R)dely[e];
        gp = e == e0 ? t : gp + t;
    }
    gp = K.iaxis == 1 ? fdiv(R(2) * gp, dhyd * dhyd) : fdiv(gp, dhyd);
    const R dnr = tsqrt(scx * scx + scy * scy);
    const R enx = fdiv(-scx, dnr), eny = fdiv(-scy, dnr);
    R rho = par.c[0];
    const R rdot = K.acomb * tpow(par.p * R(kConvKsc), K.ancomb);
    R rdoteb = rdot;
    R vel = fdiv(K.rhoprop, rho) * rdoteb;
    s.p = par.p;
    s.u = vel * enx + vel * enx - par.u;
    s.v = vel * eny + vel * eny - par.v;
    s.t = K.tflame - fdiv(R(0.5) * (s.u * s.u + s.v * s.v), K.cp);
    R amu = K.suth_sqrt_amol * tpow(s.t, R(0.6));
    const R gval = fdiv(gp, K.rhoprop * rdot) *
                   tpow(fdiv(K.rhoprop * rdot * dhyd, amu), R(-0.125));
    if (gval > R(kGth)) rdoteb = rdoteb * (R(1) + R(0.023) * (tpow(gval, R(0.8)) - K.gth_08));
    vel = fdiv(K.rhoprop, rho) * rdoteb;
    s.u = vel * enx + vel * enx - par.u;
    s.v = vel * eny + vel * eny - par.v;
    s.t = K.tflame - fdiv(R(0.5) * (s.u * s.u + s.v * s.v), K.cp);
    rho = fdiv(s.p, K.rr * s.t);
    amu = K.suth_sqrt_amol * tpow(s.t, R(0.6));
    const R rt = fdiv(amu, rho);
    ghost_complete(s, rho, rho * rt, K);
}

// One face of VISCOUS_FACE (viscous.f90:130-217).  A = cell, B/D = face nodes,
// C = neighbour.  Returns the face terms as pairs (a, b) so that the caller
// accumulates VIS = (VIS + a) + b, the reference's order.
template <class R>
struct ViscFace {
    R u0, u1, v0, v1, t0, t1, r0, r1;
};

template <class R>
__device__ __forceinline__ ViscFace<R> viscous_face(R sx, R sy, R xa, R ya, R xb, R yb, R xc,
                                                    R yc, R xd, R yd, R ua, R va, R ta, R rhoa,
                                                    R ra, R ub, R vb, R tb, R rhob, R rb, R uc,
                                                    R vc, R tc, R rhoc, R rc, R ud, R vd, R td,
                                                    R rhod, R rd, const Consts<R> &K) {
    const Recip<R> dnr((xa - xc) * (yb - yd) - (xb - xd) * (ya - yc));
    R dux = dnr.div((ua - uc) * (yb - yd) - (ub - ud) * (ya - yc));
    R dvx = dnr.div((va - vc) * (yb - yd) - (vb - vd) * (ya - yc));
    R dtx = dnr.div((ta - tc) * (yb - yd) - (tb - td) * (ya - yc));
    R drx = dnr.div((ra - rc) * (yb - yd) - (rb - rd) * (ya - yc));
    R duy = dnr.div(-((ua - uc) * (xb - xd)) + (ub - ud) * (xa - xc));
    R dvy = dnr.div(-((va - vc) * (xb - xd)) + (vb - vd) * (xa - xc));
    R dty = dnr.div(-((ta - tc) * (xb - xd)) + (tb - td) * (xa - xc));
    R dry = dnr.div(-((ra - rc) * (xb - xd)) + (rb - rd) * (xa - xc));
    R uavg = R(0.25) * (ua + ub + uc + ud);
    R vavg = R(0.25) * (va + vb + vc + vd);
    R tavg = R(0.25) * (ta + tb + tc + td);
    R rhoavg = R(0.25) * (rhoa + rhob + rhoc + rhod);
    R ravg = R(0.25) * (ra + rb + rc + rd);
    R divvel;
    if (K.iaxis == 1) {
        R yavg = R(0.5) * (yb + yd);
        R vbyy = (yavg < R(kYaxisThresh)) ? dvy : fdiv(vavg, yavg);
        divvel = dux + dvy + vbyy;
    } else {
        divvel = dux + dvy;
    }
    R amu = K.suth_sqrt_amol * tpow(tavg, R(0.6));
    R ak = fdiv(amu * K.cp, K.prandtl);
    R chi = fdiv(rhoavg * ravg, amu * R(kCmu));
    R fmu = fdiv(ttanh(R(kAlphaTurb) * chi * chi), ttanh(R(kBetaTurb) * chi * chi));
    R amut = rhoavg * ravg * fmu;
    R akt = fdiv(amut * K.cp, K.prandtlt);
    R amueff = amu + amut;
    R akeff = ak + akt;
    R txx = amueff * (dux + dux - R(kDiv23) * divvel);
    R tyy = amueff * (dvy + dvy - R(kDiv23) * divvel);
    R txy = amueff * (duy + dvx);
    R qxx = uavg * txx + vavg * txy + akeff * dtx;
    R qyy = uavg * txy + vavg * tyy + akeff * dty;
    ViscFace<R> o;
    o.u0 = txx * sx; o.u1 = txy * sy;
    o.v0 = txy * sx; o.v1 = tyy * sy;
    o.t0 = qxx * sx; o.t1 = qyy * sy;
    o.r0 = amueff * drx * sx; o.r1 = amueff * dry * sy;
    return o;
}

// Source-term inputs of one cell: the differences between the neighbours
// across faces 1 and 3 (a - c) and faces 2 and 4 (b - d) of x, y, |V|, nu~, u,
// v.  The gradients of source_terms.f90:46-60 use only these differences, so a
// cell can fold its faces in as it visits them.
template <class R>
struct SrcDiff {
    R xac, yac, qac, rac, uac, vac;
    R xbd, ybd, qbd, rbd, ubd, vbd;
};

// neighbour values across one face: x, y, |V| = sqrt(u^2 + v^2), nu~ = C5/C1, u, v
template <class R>
struct SrcNbr {
    R x, y, q, r, u, v;
};

template <class R>
__device__ __forceinline__ SrcNbr<R> src_nbr(R x, R y, R u, R v, R c1, R c5) {
    return SrcNbr<R>{x, y, tsqrt(u * u + v * v), Recip<R>(c1).div(c5), u, v};
}

template <class R>
__device__ __forceinline__ void src_pair(SrcNbr<R> a, SrcNbr<R> c, R &dx, R &dy, R &dq, R &dr,
                                         R &du, R &dv) {
    dx = a.x - c.x; dy = a.y - c.y; dq = a.q - c.q; dr = a.r - c.r; du = a.u - c.u; dv = a.v - c.v;
}

// COMPUTE_SOURCE_TERMS for one cell (source_terms.f90:40-101).  Returns
// SOURCE5 and sets src3 (0 unless IAXIS = 1).
template <class R>
__device__ __forceinline__ R source_terms_d(const SrcDiff<R> &d, R rho, R c5, R t, R p, R v,
                                            R ycc, R &src3, const Consts<R> &K) {
    const Recip<R> dnr(d.xac * d.ybd - d.xbd * d.yac);
    R dqx = dnr.div(d.qac * d.ybd - d.qbd * d.yac);
    R drx = dnr.div(d.rac * d.ybd - d.rbd * d.yac);
    R dux = dnr.div(d.uac * d.ybd - d.ubd * d.yac);
    R dvx = dnr.div(d.vac * d.ybd - d.vbd * d.yac);
    R dqy = dnr.div(-(d.qac * d.xbd) + d.qbd * d.xac);
    R dry = dnr.div(-(d.rac * d.xbd) + d.rbd * d.xac);
    R duy = dnr.div(-(d.uac * d.xbd) + d.ubd * d.xac);
    R dvy = dnr.div(-(d.vac * d.xbd) + d.vbd * d.xac);
    const Recip<R> rrho(rho);
    R ri = rrho.div(c5);
    R amu = K.suth_sqrt_amol * tpow(t, R(0.6));
    R chi = fdiv(rho * ri, R(kCmu) * amu);
    R fmu = fdiv(ttanh(R(kAlphaTurb) * chi * chi), ttanh(R(kBetaTurb) * chi * chi));
    R fr2 = fdiv(ttanh(tsqrt(chi)), ttanh(chi * K.sqcmu));
    R amut = fmu * rho * ri;
    R phi = dqx * drx + dqy * dry;
    R dee = R(0);
    if (phi > R(0)) dee = drx * drx + dry * dry;
    R ratio = ipow(fdiv(amut, dmax(amut, R(kAmuThresh) * amu)), K.npseudo);
    R s5;
    if (K.iaxis == 1) {
        const Recip<R> rycc(ycc);
        R vbyy = (ycc < R(kYaxisThresh)) ? dvy : rycc.div(v);
        R divvel = dux + dvy + vbyy;
        R pk = rrho.div(amut) *
               (dux * dux + dvy * dvy + vbyy * vbyy + dux * dux + dvy * dvy + vbyy * vbyy +
                (duy + dvx) * (duy + dvx) - R(kDiv23) * divvel * divvel);
        pk = dmax(pk, R(0));
        s5 = (rho * (R(kConst1) - R(kConst2) * fr2) * tsqrt(ri * pk) - rho * R(kConst3) * dee) *
             ratio;
        R ttt = (amu + amut) * (vbyy + vbyy - R(kDiv23) * divvel);
        src3 = rycc.div(p - ttt);
    } else {
        R divvel = dux + dvy;
        R pk = rrho.div(amut) *
               (dux * dux + dvy * dvy + dux * dux + dvy * dvy + (duy + dvx) * (duy + dvx) -
                R(kDiv23) * divvel * divvel);
        pk = dmax(pk, R(0));
        s5 = (rho * (R(kConst1) - R(kConst2) * fr2) * tsqrt(ri * pk) - rho * R(kConst3) * dee) *
             ratio;
        src3 = R(0);
    }
    return s5;
}

// The same from the four neighbours' values (a..d = across faces 1..4).
template <class R>
__device__ __forceinline__ R source_terms(const SrcNbr<R> &a, const SrcNbr<R> &b,
                                          const SrcNbr<R> &c, const SrcNbr<R> &dd, R rho, R c5,
                                          R t, R p, R v, R ycc, R &src3, const Consts<R> &K) {
    SrcDiff<R> d;
    src_pair(a, c, d.xac, d.yac, d.qac, d.rac, d.uac, d.vac);
    src_pair(b, dd, d.xbd, d.ybd, d.qbd, d.rbd, d.ubd, d.vbd);
    return source_terms_d(d, rho, c5, t, p, v, ycc, src3, K);
}

// The same with the neighbour's VOL/DT given (vbn = VOL(n)/DT(n), the very
// division jst_coeffs performs).
template <class R>
__device__ __forceinline__ void jst_coeffs_v(R si, R sn, R volbydt, R vbn, R &d2n, R &d4n) {
    R e2n = R(kC02) * dmax(si, sn);
    R e4n = dmax(R(0), R(kC04) - e2n);
    R d24n = R(0.5) * (volbydt + vbn);
    d2n = e2n * d24n;
    d4n = e4n * d24n;
}

// Dissipation face coefficients (solver_state.f90:192-194): returns D2N, D4N.
template <class R>
__device__ __forceinline__ void jst_coeffs(R si, R sn, R volbydt, R vol_n, R dt_n, R &d2n,
                                           R &d4n) {
    R e2n = R(kC02) * dmax(si, sn);
    R e4n = dmax(R(0), R(kC04) - e2n);
    R d24n = R(0.5) * (volbydt + fdiv(vol_n, dt_n));
    d2n = e2n * d24n;
    d4n = e4n * d24n;
}

}  // namespace tdt
