// Adjoint of one full iteration (turbnstwod.f90:95-145): the pieces that the
// residual adjoint (adjoint.hip) does not cover.
//
// One iteration maps Q^n to Q^{n+1} through
//   DT = DT(Q^n)                 (turbnstwod.f90:97-108, LOCAL = 1)
//   D  = D(Q^n)                  (DISSIPATION, solver_state.f90:70-225)
//   V  = VIS(Q^n)                (COMPUTE_VISCOUS at stage 1, viscous.f90:29-217)
//   Q^(0) = Q^n,  X_k = Q^n - a_k s (R(Q^(k-1)) - V - D),  Q^(k) = clamp(X_k)
// with a_k = ALPHA_k DT / VOL, s = (1, 1, 1, 1, CFLTURB) and clamp the nu~
// floor C5 = C1 max(RTOL, C5 / C1) (residual.f90:67-83).
//
// E_i = D_i + V_i (V has no mass row) and DT_i read the states of a small set
// of cells around i: the face neighbours, their face neighbours (the sensors
// and D2Q of the dissipation), and the cells and ghost parents of the node
// averages at i's four nodes.  The host lists these members per cell.
//   k_op_jac   forms, per cell and member (one thread per dual direction), the
//              6x5 block d(E_i, DT_i)/dQ_member
//              by forward-mode dual numbers through the same arithmetic as the
//              routine kernels (k_timestep, k_dissip1/2, k_prim, k_bc,
//              k_nodeavg, k_viscous), so the value part is theirs bit for bit;
//              k_op_jac_par dE_i/d(P0, T0, RTOL) and k_op_jac_xy
//              d(E_i, DT_i, VOL_i)/d(node x, y) over the residual adjoint's 20
//              node slots;
//   k_op_vjp   the transposed gather over the members (deterministic, no
//              atomics), k_op_par_vjp / k_op_par_fold the parameter sums, and
//              k_op_xy_vjp the node gather;
//   k_stage_vjp the reverse of one stage update and its clamp (elementwise).
#include "ctx.hpp"
#include "dual.hpp"
#include "reduce.hpp"

namespace tdt {
namespace {

using std::fabs;   // FP64 operands; Dual ones find tdt::fabs by argument lookup

// directions per dual evaluation: fewer directions per thread, more threads
constexpr int kDualN = 1;
constexpr int kOpRows = 6;                 // member block rows: E1..E5, DT
constexpr int kOpBlock = kOpRows * 5;      // [row][c]
constexpr int kOpParRows = 5;              // E1..E5 x (P0, T0, RTOL)
constexpr int kOpXyRows = 7;               // node rows: E1..E5, DT, VOL
constexpr int kNodeSlots = 20;
constexpr int kOpXyStride = kNodeSlots * 2 * kOpXyRows;   // [slot][x, y][row]

__device__ __forceinline__ int gid() { return blockIdx.x * kBlock + threadIdx.x; }

// ---- geometry -------------------------------------------------------------------

// The geometry cell i's operators read, from the stored arrays.
template <class R>
struct OpGeoStored {
    const Dev *D;
    int i;
    __device__ R x(int v) const { return R(D->xn[v]); }
    __device__ R y(int v) const { return R(D->yn[v]); }
    __device__ R xc(int id) const { return R(D->xcc[id]); }
    __device__ R yc(int id) const { return R(D->ycc[id]); }
    __device__ R vol(int id) const { return R(D->vol[id]); }
    __device__ R dl(int id) const { return R(D->dl[id]); }
    __device__ R sc(int k) const { return R(D->sc[k * D->n + i]); }
};

// ... recomputed from the node slots (the residual adjoint's numbering: 0-3 the
// cell's nodes 1-4, 4 + 4f + k node k+1 of the cell across face f+1) with the
// arithmetic of k_geometry and k_geometry_ghosts, so derivatives with respect
// to the slot coordinates follow.  Ghosts read here are those of the cell's own
// faces and, in node averages, of its face neighbours' faces (checked on the
// host).
template <class R>
struct OpGeoSlots {
    const Dev *D;
    int i;
    int nb[4];
    int d0;                 // first direction: direction d0 + k seeds slot (d0 + k) / 2,
                            // its x if even, its y if odd
    int vfirst[kDualN];     // the node whose lookups by id carry direction k: the
                            // seeded slot's node if that slot is its first, else -1
    R xcc[5], ycc[5], volc[5], dlc[5];     // the cell, then across faces 1..4
    R scv[8];

    __device__ int node_of(int s) const {
        const int *nod = reinterpret_cast<const int *>(D->nod4);
        if (s < 4) return nod[4 * i + s];
        const int j = nb[(s - 4) >> 2];
        return j < D->n ? nod[4 * j + ((s - 4) & 3)] : -1;
    }
    // the coordinates of slot s, seeded (read from the node arrays on demand:
    // kept in arrays, their lookups by id would index them dynamically)
    __device__ R xs(int s) const {
        const int v = node_of(s);
        R r(v >= 0 ? D->xn[v] : 0.0);
        for (int k = 0; k < kDualN; ++k)
            if (d0 + k == 2 * s) r.d[k] = 1.0;
        return r;
    }
    __device__ R ys(int s) const {
        const int v = node_of(s);
        R r(v >= 0 ? D->yn[v] : 0.0);
        for (int k = 0; k < kDualN; ++k)
            if (d0 + k == 2 * s + 1) r.d[k] = 1.0;
        return r;
    }
    __device__ void seed(int first_dir) {
        d0 = first_dir;
        for (int k = 0; k < kDualN; ++k) {
            vfirst[k] = -1;
            const int sl = (d0 + k) >> 1;
            if (d0 + k >= 2 * kNodeSlots) continue;
            const int v = node_of(sl);
            int first = sl;
            for (int t = 0; t < sl; ++t)
                if (node_of(t) == v) {
                    first = t;
                    break;
                }
            if (first == sl) vfirst[k] = v;
        }
    }
    __device__ int face_of(int id) const {
        for (int f = 0; f < 4; ++f)
            if (nb[f] == id) return f;
        return -1;
    }
    // a node's coordinates by id: those of the first slot holding it
    __device__ R x(int v) const {
        R r(D->xn[v]);
        for (int k = 0; k < kDualN; ++k)
            if (vfirst[k] == v && ((d0 + k) & 1) == 0) r.d[k] = 1.0;
        return r;
    }
    __device__ R y(int v) const {
        R r(D->yn[v]);
        for (int k = 0; k < kDualN; ++k)
            if (vfirst[k] == v && ((d0 + k) & 1) == 1) r.d[k] = 1.0;
        return r;
    }
    __device__ void centroid(int id, R &cx, R &cy) const {
        if (id == i) { cx = xcc[0]; cy = ycc[0]; return; }
        const int f = face_of(id);
        if (f >= 0) { cx = xcc[1 + f]; cy = ycc[1 + f]; return; }
        // a ghost of a face neighbour j: mirrored through j's face midpoint
        const int k = id - D->n, par = D->gparent[k], sd = D->gside[k];
        const int fj = face_of(par);
        const int b = 4 + 4 * fj;
        const int a = b + sd - 1, c = b + (sd & 3);
        const R xf = R(0.5) * (xs(a) + xs(c));
        const R yf = R(0.5) * (ys(a) + ys(c));
        cx = R(2.0) * xf - xcc[1 + fj];
        cy = R(2.0) * yf - ycc[1 + fj];
    }
    __device__ R xc(int id) const { R a, b; centroid(id, a, b); return a; }
    __device__ R yc(int id) const { R a, b; centroid(id, a, b); return b; }
    __device__ R vol(int id) const { return id == i ? volc[0] : volc[1 + face_of(id)]; }
    __device__ R dl(int id) const { return id == i ? dlc[0] : dlc[1 + face_of(id)]; }
    __device__ R sc(int k) const { return scv[k]; }
};

// centroid, min edge and volume of a quad from its nodes (k_geometry)
template <class R>
__device__ void quad_geo(const R (&x)[4], const R (&y)[4], int iaxis, R &xc, R &yc, R &dl,
                         R &vol, R *sc) {
    xc = (x[0] + x[1] + x[2] + x[3]) * R(0.25);
    yc = (y[0] + y[1] + y[2] + y[3]) * R(0.25);
    const R dl12 = tsqrt((x[0] - x[1]) * (x[0] - x[1]) + (y[0] - y[1]) * (y[0] - y[1]));
    const R dl23 = tsqrt((x[1] - x[2]) * (x[1] - x[2]) + (y[1] - y[2]) * (y[1] - y[2]));
    const R dl34 = tsqrt((x[2] - x[3]) * (x[2] - x[3]) + (y[2] - y[3]) * (y[2] - y[3]));
    const R dl41 = tsqrt((x[3] - x[0]) * (x[3] - x[0]) + (y[3] - y[0]) * (y[3] - y[0]));
    dl = dmin(dmin(dmin(dl12, dl23), dl34), dl41);
    if (iaxis == 1) {
        if (sc) {
            sc[0] = (y[1] - y[0]) * (y[0] + y[1]) * R(0.5);
            sc[1] = -((x[1] - x[0]) * (y[0] + y[1]) * R(0.5));
            sc[2] = (y[2] - y[1]) * (y[1] + y[2]) * R(0.5);
            sc[3] = -((x[2] - x[1]) * (y[1] + y[2]) * R(0.5));
            sc[4] = (y[3] - y[2]) * (y[2] + y[3]) * R(0.5);
            sc[5] = -((x[3] - x[2]) * (y[2] + y[3]) * R(0.5));
            sc[6] = (y[0] - y[3]) * (y[3] + y[0]) * R(0.5);
            sc[7] = -((x[0] - x[3]) * (y[3] + y[0]) * R(0.5));
        }
        vol = R(0.5) * ((x[2] - x[0]) * (y[3] - y[1]) - (x[3] - x[1]) * (y[2] - y[0])) * yc;
    } else {
        if (sc) {
            sc[0] = (y[1] - y[0]);
            sc[1] = -(x[1] - x[0]);
            sc[2] = (y[2] - y[1]);
            sc[3] = -(x[2] - x[1]);
            sc[4] = (y[3] - y[2]);
            sc[5] = -(x[3] - x[2]);
            sc[6] = (y[0] - y[3]);
            sc[7] = -(x[0] - x[3]);
        }
        vol = R(0.5) * ((x[2] - x[0]) * (y[3] - y[1]) - (x[3] - x[1]) * (y[2] - y[0]));
    }
    if (vol <= R(0)) vol = -vol;
}

template <class R>
__device__ void slots_finish(OpGeoSlots<R> &G, int iaxis) {
    {
        const R x[4] = {G.xs(0), G.xs(1), G.xs(2), G.xs(3)};
        const R y[4] = {G.ys(0), G.ys(1), G.ys(2), G.ys(3)};
        quad_geo(x, y, iaxis, G.xcc[0], G.ycc[0], G.dlc[0], G.volc[0], G.scv);
    }
    for (int f = 0; f < 4; ++f) {
        if (G.nb[f] < G.D->n) {
            const int b = 4 + 4 * f;
            const R x[4] = {G.xs(b), G.xs(b + 1), G.xs(b + 2), G.xs(b + 3)};
            const R y[4] = {G.ys(b), G.ys(b + 1), G.ys(b + 2), G.ys(b + 3)};
            quad_geo(x, y, iaxis, G.xcc[1 + f], G.ycc[1 + f], G.dlc[1 + f], G.volc[1 + f],
                     (R *)nullptr);
        } else {   // k_geometry_ghosts: mirrored centroid, VOL(NG) = VOL(NP)
            const int a = f, c = (f + 1) & 3;
            const R xf = R(0.5) * (G.xs(a) + G.xs(c));
            const R yf = R(0.5) * (G.ys(a) + G.ys(c));
            G.xcc[1 + f] = R(2.0) * xf - G.xcc[0];
            G.ycc[1 + f] = R(2.0) * yf - G.ycc[0];
            G.volc[1 + f] = G.volc[0];
            G.dlc[1 + f] = G.dlc[0];
        }
    }
}

// ---- the operators of one cell -----------------------------------------------------

template <class R>
struct OpOut {
    R d[5];      // D241..D245
    R vis[4];    // VISU, VISV, VIST, VISR
    R dt, vol;
};

// k_timestep's P and DT of a state
template <class R>
__device__ __forceinline__ void tstep(const R (&q)[5], R dl, const Consts<R> &K, R &p, R &dt) {
    const R rho = q[0];
    const R u = q[1] / rho;
    const R v = q[2] / rho;
    const R vel = tsqrt(u * u + v * v);
    const R e = q[3] / rho;
    const R t = (e - R(0.5) * (u * u + v * v)) / K.cv;
    p = rho * K.rr * t;
    dt = K.cfl * dl / (fabs(vel) + tsqrt(K.gamma_rr * t));
}

template <class R>
__device__ __forceinline__ R tstep_p(const R (&q)[5], const Consts<R> &K) {
    const R rho = q[0];
    const R u = q[1] / rho;
    const R v = q[2] / rho;
    const R e = q[3] / rho;
    const R t = (e - R(0.5) * (u * u + v * v)) / K.cv;
    return rho * K.rr * t;
}

// primitives of a cell (k_prim) or of a ghost after k_bc
template <class R, class QL, class GEO>
__device__ void prim_of(const Dev &D, int id, const QL &ql, const GEO &G, const Consts<R> &K,
                        Prim<R> &s) {
    const int n = D.n;
    if (id < n) {
        R q[5];
        ql(id, q);
        prim_from_q<R>(s, q[0], q[1], q[2], q[3], q[4], K);
        return;
    }
    const int k = id - n;
    const int par = D.gparent[k], ty = D.gtype[k];
    R q[5];
    ql(par, q);
    Prim<R> pp;
    prim_from_q<R>(pp, q[0], q[1], q[2], q[3], q[4], K);
    R uo = R(0), vo = R(0), to = R(0), po = R(0);
    if (ty == 3) {
        Prim<R> po_;
        R qo[5];
        ql(D.gopp[k], qo);
        prim_from_q<R>(po_, qo[0], qo[1], qo[2], qo[3], qo[4], K);
        uo = po_.u; vo = po_.v; to = po_.t; po = po_.p;
    }
    ghost_state<R>(s, ty, pp, uo, vo, to, po, G.yc(id), K);
}

// k_timestep's C, P and DT of cell x and DISSIPATION's sensors ANU, ANURT and
// second differences D2Q there (a ghost neighbour of x counts as a copy of x)
template <class R, class QL, class GEO>
__device__ __forceinline__ void cell_sensors(const Dev &D, int x, const QL &ql, const GEO &G,
                                             const Consts<R> &K, R (&cx)[5], R &dt, R &anu,
                                             R &anurt, R (&d2q)[5]) {
    const int n = D.n;
    R px;
    ql(x, cx);
    tstep<R>(cx, G.dl(x), K, px, dt);
    const int4 y4 = D.nc4[x];
    const int ys[4] = {y4.x, y4.y, y4.z, y4.w};
    R an, ad, bn, bd, sum[5];
    for (int k = 0; k < 4; ++k) {
        R cy[5], py;
        if (ys[k] < n) {
            ql(ys[k], cy);
            py = tstep_p<R>(cy, K);
        } else {
            for (int m = 0; m < 5; ++m) cy[m] = cx[m];
            py = px;
        }
        const R tn = fabs(py - px), td = fabs(py + px);
        const R un = fabs(cy[4] - cx[4]), ud = fabs(cy[4] + cx[4]);
        if (k == 0) {
            an = tn; ad = td; bn = un; bd = ud;
            for (int m = 0; m < 5; ++m) sum[m] = cy[m];
        } else {
            an = an + tn; ad = ad + td; bn = bn + un; bd = bd + ud;
            for (int m = 0; m < 5; ++m) sum[m] = sum[m] + cy[m];
        }
    }
    anu = an / ad;
    anurt = bn / dmax(bd, R(8.0) * cx[0] * K.rtol);
    for (int m = 0; m < 5; ++m) d2q[m] = sum[m] - R(4.0) * cx[m];
}

// k_nodeavg's UN, VN, TN, RHON, RN at node v (the cells and ghosts around it,
// in the node list's order)
template <class R, class QL, class GEO>
__device__ __forceinline__ void node_avg(const Dev &D, int v, const QL &ql, const GEO &G,
                                         const Consts<R> &K, R (&a)[5]) {
    const int4 ov = D.novr[v];
    const int j0 = D.nptr[v], j1 = D.nptr[v + 1];
    R au = R(0), av = R(0), at = R(0), arho = R(0);
    R ar = ov.w ? K.rtol_rtred : R(0);
    for (int j = j0; j < j1; ++j) {
        Prim<R> s;
        prim_of<R>(D, D.nlist[j], ql, G, K, s);
        const R rho = s.c[0], rr = s.c[4] / rho;
        if (j >= ov.x) au = au + s.u;
        if (j >= ov.y) av = av + s.v;
        at = at + s.t;
        arho = arho + rho;
        if (j >= ov.z) ar = ar + rr;
    }
    const R cnt = R((double)(j1 - j0));
    a[0] = au / cnt;
    a[1] = av / cnt;
    a[2] = at / cnt;
    a[3] = arho / cnt;
    a[4] = dmax(K.rtol, ar / cnt);
}

template <class R, class QL, class GEO>
__device__ void cell_ops(const Dev &D, int i, const QL &ql, const GEO &G, const Consts<R> &K,
                         OpOut<R> &o) {
    const int n = D.n;
    const int4 nb4 = D.nc4[i];
    const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
    // --- DISSIPATION with the timestep's P and DT: the cell's C, DT, sensors and
    // D2Q, then each face neighbour's (a ghost: the cell's, ANURT 0), whose face
    // terms are added as soon as they are known (each D2/D4 sum still runs over
    // the faces in order), so only one neighbour's values are live at a time
    R c0[5], d2q0[5], dt0, anu0, anurt0;
    cell_sensors<R>(D, i, ql, G, K, c0, dt0, anu0, anurt0, d2q0);
    const R vol0 = G.vol(i);
    const R volbydt = vol0 / dt0;
    R d2[5], d4[5];
    for (int k = 0; k < 4; ++k) {
        const int x = nb[k];
        R cx[5], d2qx[5], dtx, anux, anurtx, volx;
        if (x < n) {
            cell_sensors<R>(D, x, ql, G, K, cx, dtx, anux, anurtx, d2qx);
            volx = G.vol(x);
        } else {
            for (int m = 0; m < 5; ++m) {
                cx[m] = c0[m];
                d2qx[m] = d2q0[m];
            }
            dtx = dt0;
            anux = anu0;
            anurtx = R(0);
            volx = vol0;
        }
        R a2, a4, b2, b4;
        jst_coeffs<R>(anu0, anux, volbydt, volx, dtx, a2, a4);
        jst_coeffs<R>(anurt0, anurtx, volbydt, volx, dtx, b2, b4);
        for (int m = 0; m < 5; ++m) {
            const R t2 = (m < 4 ? a2 : b2) * (cx[m] - c0[m]);
            const R t4 = (m < 4 ? a4 : b4) * (d2qx[m] - d2q0[m]);
            d2[m] = k == 0 ? t2 : d2[m] + t2;
            d4[m] = k == 0 ? t4 : d4[m] + t4;
        }
    }
    for (int m = 0; m < 5; ++m) o.d[m] = d2[m] - d4[m];
    o.dt = dt0;
    o.vol = vol0;
    // --- COMPUTE_VISCOUS at stage 1: the faces in order, each between nodes f
    // and f + 1, whose averages are formed as the faces reach them (node 1's is
    // kept for face 4)
    const int4 nd4 = D.nod4[i];
    const int nd[4] = {nd4.x, nd4.y, nd4.z, nd4.w};
    Prim<R> pi;
    prim_from_q<R>(pi, c0[0], c0[1], c0[2], c0[3], c0[4], K);
    const R xa = G.xc(i), ya = G.yc(i);
    const R rhoa = pi.c[0], ra = pi.c[4] / rhoa;
    R n1[5], B[5];
    node_avg<R>(D, nd[0], ql, G, K, n1);
    for (int m = 0; m < 5; ++m) B[m] = n1[m];
    R vu = R(0), vv = R(0), vt = R(0), vr = R(0);
    for (int f = 0; f < 4; ++f) {
        const int b = nd[f], d = nd[(f + 1) & 3], c = nb[f];
        R Dn[5];
        if (f < 3) {
            node_avg<R>(D, d, ql, G, K, Dn);
        } else {
            for (int m = 0; m < 5; ++m) Dn[m] = n1[m];
        }
        Prim<R> pc;
        prim_of<R>(D, c, ql, G, K, pc);
        const R rhoc = pc.c[0];
        ViscFace<R> w = viscous_face<R>(G.sc(2 * f), G.sc(2 * f + 1), xa, ya, G.x(b), G.y(b),
                                        G.xc(c), G.yc(c), G.x(d), G.y(d), pi.u, pi.v, pi.t, rhoa,
                                        ra, B[0], B[1], B[2], B[3], B[4], pc.u, pc.v, pc.t, rhoc,
                                        pc.c[4] / rhoc, Dn[0], Dn[1], Dn[2], Dn[3], Dn[4], K);
        vu = vu + w.u0 + w.u1;
        vv = vv + w.v0 + w.v1;
        vt = vt + w.t0 + w.t1;
        vr = vr + w.r0 + w.r1;
        for (int m = 0; m < 5; ++m) B[m] = Dn[m];
    }
    o.vis[0] = vu; o.vis[1] = vv; o.vis[2] = vt; o.vis[3] = vr;
}

// ---- state loaders --------------------------------------------------------------------

struct ValueLoader {
    const double *qs;
    int n;
    __device__ void operator()(int id, double (&q)[5]) const {
        for (int c = 0; c < 5; ++c) q[c] = qs[(size_t)c * n + id];
    }
};

// member `seed` carries directions k = 0..kDualN-1: dQ_seed,(c0 + k)
struct SeedLoader {
    const double *qs;
    int n, seed, c0;
    __device__ void operator()(int id, Dual<kDualN> (&q)[5]) const {
        for (int c = 0; c < 5; ++c) {
            q[c] = Dual<kDualN>(qs[(size_t)c * n + id]);
            if (id == seed && c >= c0 && c < c0 + kDualN) q[c].d[c - c0] = 1.0;
        }
    }
};

struct ConstLoader {
    const double *qs;
    int n;
    __device__ void operator()(int id, Dual<kDualN> (&q)[5]) const {
        for (int c = 0; c < 5; ++c) q[c] = Dual<kDualN>(qs[(size_t)c * n + id]);
    }
};

// ---- kernels -----------------------------------------------------------------------

// D, VIS, DT of the state qs with the routine kernels' arithmetic (checks the
// evaluation against them): od = [D1..D5][n], ov = [VIS1..4, DT][n]
__global__ void __launch_bounds__(kBlock) k_op_eval(Dev D, Consts<double> K, const double *qs,
                                                    double *od, double *ov) {
    const int i = gid();
    if (i >= D.n) return;
    OpGeoStored<double> G{&D, i};
    OpOut<double> o;
    cell_ops<double>(D, i, ValueLoader{qs, D.n}, G, K, o);
    const size_t n = D.n;
    for (int m = 0; m < 5; ++m) od[m * n + i] = o.d[m];
    for (int m = 0; m < 4; ++m) ov[m * n + i] = o.vis[m];
    ov[4 * n + i] = o.dt;
}

struct OpSeeds {
    double ds0_dp0, ds0_dt0, dtinf_dt0, dblu_dt0, rtred;
};

constexpr int kMemberGroups = (5 + kDualN - 1) / kDualN;             // per member
constexpr int kParGroups = (3 + kDualN - 1) / kDualN;                // P0, T0, RTOL
constexpr int kNodeGroups = (2 * kNodeSlots + kDualN - 1) / kDualN;  // slot x, y

// One dual evaluation per thread, in three kernels (each branch gets the
// registers it needs alone).  k_op_jac, blockIdx.y = e: for cell i, member m = e
// / kMemberGroups (m < its member count) with components c0.. of that member
// seeded: JE[mptr[i] + m] = d(E_i, DT_i)/dQ_member [row][c].  k_op_jac_par,
// blockIdx.y = e: parameter group e, JEp[i] = dE_i/d(P0, T0, RTOL) [row][p].
// k_op_jac_xy, blockIdx.y = e: node group e, JEx[i] = d(E_i, DT_i, VOL_i)/d(slot
// x, y).
__global__ void __launch_bounds__(64, 2) k_op_jac(Dev D, Consts<double> K, const double *qs,
                                               const int *mptr, const int *mlist, double *JE) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int e = blockIdx.y;
    if (i >= D.n) return;
    const int n = D.n;
    const Consts<Dual<kDualN>> KD = lift<kDualN>(K);
    const int ent = mptr[i] + e / kMemberGroups;
    const int c0 = (e % kMemberGroups) * kDualN;
    if (ent >= mptr[i + 1]) return;
    OpGeoStored<Dual<kDualN>> G{&D, i};
    OpOut<Dual<kDualN>> o;
    cell_ops<Dual<kDualN>>(D, i, SeedLoader{qs, n, mlist[ent], c0}, G, KD, o);
    double *B = JE + (size_t)ent * kOpBlock;
    for (int k = 0; k < kDualN && c0 + k < 5; ++k) {
        const int c = c0 + k;
        B[c] = o.d[0].d[k];
        for (int m = 1; m < 5; ++m) B[m * 5 + c] = o.d[m].d[k] + o.vis[m - 1].d[k];
        B[25 + c] = o.dt.d[k];
    }
}

__global__ void __launch_bounds__(64) k_op_jac_par(Dev D, Consts<double> K, OpSeeds ps,
                                                   const double *qs, double *JEp) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int e = blockIdx.y;
    if (i >= D.n) return;
    const int n = D.n;
    const Consts<Dual<kDualN>> KD = lift<kDualN>(K);
    // parameter p = e * kDualN + k in direction k
    Consts<Dual<kDualN>> KP = KD;
    for (int k = 0; k < kDualN; ++k) {
        const int p = e * kDualN + k;
        if (p == 0) KP.s0.d[k] = ps.ds0_dp0;
        if (p == 1) {
            KP.s0.d[k] = ps.ds0_dt0;
            KP.t0.d[k] = 1.0;
            KP.tinf.d[k] = ps.dtinf_dt0;
            KP.bl_u.d[k] = ps.dblu_dt0;
        }
        if (p == 2) {
            KP.rtol.d[k] = 1.0;
            KP.rtol_rtred.d[k] = ps.rtred;
        }
    }
    OpGeoStored<Dual<kDualN>> G{&D, i};
    OpOut<Dual<kDualN>> o;
    cell_ops<Dual<kDualN>>(D, i, ConstLoader{qs, n}, G, KP, o);
    double *P = JEp + (size_t)i * kOpParRows * 3;
    for (int k = 0; k < kDualN && e * kDualN + k < 3; ++k) {
        const int p = e * kDualN + k;
        P[p] = o.d[0].d[k];
        for (int m = 1; m < 5; ++m) P[m * 3 + p] = o.d[m].d[k] + o.vis[m - 1].d[k];
    }
}

__global__ void __launch_bounds__(64) k_op_jac_xy(Dev D, Consts<double> K, const double *qs,
                                                  double *JEx) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int e = blockIdx.y;
    if (i >= D.n) return;
    const int n = D.n;
    if (e >= kNodeGroups) return;
    const Consts<Dual<kDualN>> KD = lift<kDualN>(K);
    const int d0 = e * kDualN;   // first of this group's directions
    OpGeoSlots<Dual<kDualN>> GS;
    GS.D = &D;
    GS.i = i;
    const int4 nb4 = D.nc4[i];
    GS.nb[0] = nb4.x; GS.nb[1] = nb4.y; GS.nb[2] = nb4.z; GS.nb[3] = nb4.w;
    GS.seed(d0);
    slots_finish(GS, K.iaxis);
    OpOut<Dual<kDualN>> o;
    cell_ops<Dual<kDualN>>(D, i, ConstLoader{qs, n}, GS, KD, o);
    double *X = JEx + (size_t)i * kOpXyStride;
    for (int k = 0; k < kDualN && d0 + k < 2 * kNodeSlots; ++k) {
        double *Xr = X + (size_t)(d0 + k) * kOpXyRows;
        Xr[0] = o.d[0].d[k];
        for (int m = 1; m < 5; ++m) Xr[m] = o.d[m].d[k] + o.vis[m - 1].d[k];
        Xr[5] = o.dt.d[k];
        Xr[6] = o.vol.d[k];
    }
}

// g_k = sum over the entries (i, e) of cell k's transposed member list, in
// order, of w_i^T JE[e], w_i = (W_i, gDT_i)
__global__ void __launch_bounds__(kBlock) k_op_vjp(Dev D, const double *JE, const double *W,
                                                   const double *gdt, const int *tptr,
                                                   const int2 *tlist, double *g) {
    const int k = gid();
    if (k >= D.n) return;
    const int n = D.n;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int e = tptr[k]; e < tptr[k + 1]; ++e) {
        const int2 ie = tlist[e];
        const double *B = JE + (size_t)ie.y * kOpBlock;
        double w[kOpRows];
        for (int m = 0; m < 5; ++m) w[m] = W[(size_t)m * n + ie.x];
        w[5] = gdt[ie.x];
        for (int c = 0; c < 5; ++c) {
            double t = w[0] * B[c];
            for (int m = 1; m < kOpRows; ++m) t = t + w[m] * B[m * 5 + c];
            acc[c] = acc[c] + t;
        }
    }
    for (int c = 0; c < 5; ++c) g[(size_t)c * n + k] = acc[c];
}

// per block: sums over its cells of W_i^T dE_i/d(P0, T0, RTOL), and of the clamp's
// RTOL term grtol_i
__global__ void __launch_bounds__(kBlock) k_op_par_vjp(Dev D, const double *JEp, const double *W,
                                                       const double *grtol, double *part) {
    const int i = gid();
    double t[3] = {0.0, 0.0, 0.0};
    if (i < D.n) {
        const double *P = JEp + (size_t)i * kOpParRows * 3;
        for (int p = 0; p < 3; ++p) {
            double a = W[i] * P[p];
            for (int m = 1; m < 5; ++m) a = a + W[(size_t)m * D.n + i] * P[m * 3 + p];
            t[p] = a;
        }
        t[2] = t[2] + grtol[i];
    }
    __shared__ double sh[kBlock / kWave][3];
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    for (int p = 0; p < 3; ++p) {
        const double s = wave_sum(t[p]);
        if (lane == 0) sh[w][p] = s;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double a = sh[0][threadIdx.x];
        for (int q = 1; q < kBlock / kWave; ++q) a = a + sh[q][threadIdx.x];
        part[blockIdx.x * 3 + threadIdx.x] = a;
    }
}

// out[p] += block partials in block order
__global__ void k_op_par_fold(const double *part, int nblk, double *out) {
    const int p = threadIdx.x;
    if (p >= 3) return;
    double a = 0.0;
    for (int b = 0; b < nblk; ++b) a = a + part[b * 3 + p];
    out[p] = out[p] + a;
}

// gxy_v += sum over the (cell i, slot s) whose slot holds node v, in order, of
// (W_i, gDT_i, gVOL_i)^T d(E_i, DT_i, VOL_i)/d(x, y)_s
__global__ void __launch_bounds__(kBlock) k_op_xy_vjp(Dev D, const double *JEx, const double *W,
                                                      const double *gdt, const double *gvol,
                                                      const int *nptr, const int2 *nlist,
                                                      double *gxy) {
    const int nd = gid();
    if (nd >= D.m) return;
    const int n = D.n;
    double ax = 0.0, ay = 0.0;
    for (int e = nptr[nd]; e < nptr[nd + 1]; ++e) {
        const int2 is = nlist[e];
        const double *B = JEx + (size_t)is.x * kOpXyStride + (size_t)is.y * 2 * kOpXyRows;
        double w[kOpXyRows];
        for (int m = 0; m < 5; ++m) w[m] = W[(size_t)m * n + is.x];
        w[5] = gdt[is.x];
        w[6] = gvol[is.x];
        double tx = w[0] * B[0], ty = w[0] * B[kOpXyRows];
        for (int r = 1; r < kOpXyRows; ++r) {
            tx = tx + w[r] * B[r];
            ty = ty + w[r] * B[kOpXyRows + r];
        }
        ax = ax + tx;
        ay = ay + ty;
    }
    gxy[nd] = gxy[nd] + ax;
    gxy[D.m + nd] = gxy[D.m + nd] + ay;
}

// Reverse of one stage update (k_update) given mu = adjoint of Q^(k):
//   lam += xi (the COLD term), wr = -a s xi (the weight of R(Q^(k-1))),
//   W += a s xi (D and VIS), gdt / gvol += d/dDT, d/dVOL, grtol += the clamp's
// RTOL term, with xi the adjoint of the unclamped X_k.  q, r, vis, d24, dt and
// vol are Q^n, R(Q^(k-1)), V, D, DT and VOL (the cells' [comp][n] layout).
__global__ void __launch_bounds__(kBlock) k_stage_vjp(int n, double alpha, double rtol,
                                                      const double *q, const double *r,
                                                      const double *vis, const double *d24,
                                                      const double *dt, const double *vol,
                                                      const double *mu, double *lam, double *wr,
                                                      double *W, double *gdt, double *gvol,
                                                      double *grtol) {
    const int i = gid();
    if (i >= n) return;
    const double a = alpha * (dt[i] / vol[i]);
    double res[5];
    res[0] = r[i] - d24[i];
    for (int m = 1; m < 5; ++m)
        res[m] = r[(size_t)m * n + i] - vis[(size_t)(m - 1) * n + i] - d24[(size_t)m * n + i];
    const double x1 = q[i] - a * res[0];
    const double x5 = q[(size_t)4 * n + i] - a * kCflTurb * res[4];
    double xi[5];
    for (int m = 0; m < 5; ++m) xi[m] = mu[(size_t)m * n + i];
    // C5 = C1 max(RTOL, C5/C1): on the RTOL branch C5 = C1 RTOL; otherwise
    // C5 = C1 (C5/C1), whose derivatives are 0 in C1 and 1 in C5
    if (rtol > x5 / x1) {
        xi[0] = xi[0] + xi[4] * rtol;
        grtol[i] = grtol[i] + xi[4] * x1;
        xi[4] = 0.0;
    }
    double sdt = 0.0;
    for (int m = 0; m < 5; ++m) {
        const double s = m == 4 ? kCflTurb : 1.0;
        const double w = a * s * xi[m];
        lam[(size_t)m * n + i] = lam[(size_t)m * n + i] + xi[m];
        wr[(size_t)m * n + i] = -w;
        W[(size_t)m * n + i] = W[(size_t)m * n + i] + w;
        sdt = sdt + xi[m] * s * res[m];
    }
    // X = Q - alpha (DT / VOL) s res
    gdt[i] = gdt[i] - alpha / vol[i] * sdt;
    gvol[i] = gvol[i] + alpha * dt[i] / (vol[i] * vol[i]) * sdt;
}

// y += x over count values
__global__ void __launch_bounds__(kBlock) k_add(int count, const double *x, double *y) {
    const int i = gid();
    if (i < count) y[i] = y[i] + x[i];
}

__global__ void __launch_bounds__(kBlock) k_zero(int count, double *p) {
    const int i = gid();
    if (i < count) p[i] = 0.0;
}

// the fixed-point iteration's two clears of [5][n] values and one of [n], in one launch
__global__ void __launch_bounds__(kBlock) k_zero_5n_5n_n(int n, double *a, double *b, double *c) {
    const int i = gid();
    if (i < 5 * n) {
        a[i] = 0.0;
        b[i] = 0.0;
    }
    if (i < n) c[i] = 0.0;
}

// y = ((y + x1) + x2) + x3: three k_add launches in one, the same sums in the same order
__global__ void __launch_bounds__(kBlock) k_add3(int count, const double *x1, const double *x2,
                                                 const double *x3, double *y) {
    const int i = gid();
    if (i < count) y[i] = y[i] + x1[i] + x2[i] + x3[i];
}

// cells' [comp][nt] rows -> [comp][n] (and back)
__global__ void __launch_bounds__(kBlock) k_pack(int n, int nt, const double *src, double *dst) {
    const int i = gid();
    if (i < n)
        for (int m = 0; m < 5; ++m) dst[(size_t)m * n + i] = src[(size_t)m * nt + i];
}

__global__ void __launch_bounds__(kBlock) k_unpack(int n, int nt, const double *src,
                                                   double *dst, double *cold) {
    const int i = gid();
    if (i < n)
        for (int m = 0; m < 5; ++m) {
            dst[(size_t)m * nt + i] = src[(size_t)m * n + i];
            cold[(size_t)m * n + i] = src[(size_t)m * n + i];
        }
}

}  // namespace

hipError_t launch_op_eval(tdt_ctx *c, const double *qs, double *od, double *ov) {
    hipLaunchKernelGGL(k_op_eval, dim3(nblocks(c->D.n)), dim3(kBlock), 0, c->stream, c->D, c->K,
                       qs, od, ov);
    return hipGetLastError();
}

hipError_t launch_op_jac(tdt_ctx *c, const double *qs, const int *mptr, const int *mlist,
                         int kmax, double *JE, double *JEp, double *JEx) {
    const Consts<double> &K = c->K;
    const tdt_params &P = c->params;
    OpSeeds ps;
    const double gam = P.gama;
    ps.ds0_dp0 = (1.0 - gam) * (double)K.s0 / P.p0;
    ps.ds0_dt0 = gam * (double)K.s0 / P.t0;
    ps.dtinf_dt0 = (double)K.tinf / P.t0;
    ps.dblu_dt0 = 0.5 * (double)K.bl_u / P.t0;
    ps.rtred = c->params_rtred;
    const unsigned nb = (c->D.n + 63) / 64;
    hipLaunchKernelGGL(k_op_jac, dim3(nb, kmax * kMemberGroups), dim3(64), 0, c->stream, c->D, K,
                       qs, mptr, mlist, JE);
    hipLaunchKernelGGL(k_op_jac_par, dim3(nb, kParGroups), dim3(64), 0, c->stream, c->D, K, ps,
                       qs, JEp);
    if (JEx)
        hipLaunchKernelGGL(k_op_jac_xy, dim3(nb, kNodeGroups), dim3(64), 0, c->stream, c->D, K,
                           qs, JEx);
    return hipGetLastError();
}

hipError_t launch_op_vjp(tdt_ctx *c, const double *JE, const double *JEp, const double *JEx,
                         const double *W, const double *gdt, const double *gvol,
                         const double *grtol, const int *tptr, const int2 *tlist,
                         const int *nptr, const int2 *nlist, double *part, double *g,
                         double *gp, double *gxy) {
    const unsigned nb = nblocks(c->D.n);
    hipLaunchKernelGGL(k_op_vjp, dim3(nb), dim3(kBlock), 0, c->stream, c->D, JE, W, gdt, tptr,
                       tlist, g);
    if (gp) {   // gp == nullptr: the state gradient only
        hipLaunchKernelGGL(k_op_par_vjp, dim3(nb), dim3(kBlock), 0, c->stream, c->D, JEp, W,
                           grtol, part);
        hipLaunchKernelGGL(k_op_par_fold, dim3(1), dim3(kWave), 0, c->stream, part, (int)nb, gp);
    }
    if (gxy && JEx)
        hipLaunchKernelGGL(k_op_xy_vjp, dim3(nblocks(c->D.m)), dim3(kBlock), 0, c->stream, c->D,
                           JEx, W, gdt, gvol, nptr, nlist, gxy);
    return hipGetLastError();
}

hipError_t launch_stage_vjp(tdt_ctx *c, double alpha, const double *q, const double *r,
                            const double *vis, const double *d24, const double *dt,
                            const double *vol, const double *mu, double *lam, double *wr,
                            double *W, double *gdt, double *gvol, double *grtol) {
    hipLaunchKernelGGL(k_stage_vjp, dim3(nblocks(c->D.n)), dim3(kBlock), 0, c->stream, c->D.n,
                       alpha, (double)c->K.rtol, q, r, vis, d24, dt, vol, mu, lam, wr, W, gdt,
                       gvol, grtol);
    return hipGetLastError();
}

hipError_t launch_add(tdt_ctx *c, int count, const double *x, double *y) {
    hipLaunchKernelGGL(k_add, dim3(nblocks(count)), dim3(kBlock), 0, c->stream, count, x, y);
    return hipGetLastError();
}

hipError_t launch_zero(tdt_ctx *c, int count, double *p) {
    hipLaunchKernelGGL(k_zero, dim3(nblocks(count)), dim3(kBlock), 0, c->stream, count, p);
    return hipGetLastError();
}

hipError_t launch_zero_5n_5n_n(tdt_ctx *c, double *a, double *b, double *d) {
    const int n = c->D.n;
    hipLaunchKernelGGL(k_zero_5n_5n_n, dim3(nblocks(5 * n)), dim3(kBlock), 0, c->stream, n, a, b,
                       d);
    return hipGetLastError();
}

hipError_t launch_add3(tdt_ctx *c, int count, const double *x1, const double *x2,
                       const double *x3, double *y) {
    hipLaunchKernelGGL(k_add3, dim3(nblocks(count)), dim3(kBlock), 0, c->stream, count, x1, x2,
                       x3, y);
    return hipGetLastError();
}

hipError_t launch_pack(tdt_ctx *c, const double *src, double *dst) {
    hipLaunchKernelGGL(k_pack, dim3(nblocks(c->D.n)), dim3(kBlock), 0, c->stream, c->D.n,
                       c->D.nt, src, dst);
    return hipGetLastError();
}

hipError_t launch_unpack(tdt_ctx *c, const double *src) {
    hipLaunchKernelGGL(k_unpack, dim3(nblocks(c->D.n)), dim3(kBlock), 0, c->stream, c->D.n,
                       c->D.nt, src, c->D.c, c->D.cold);
    return hipGetLastError();
}

}  // namespace tdt
