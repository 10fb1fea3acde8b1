// Adjoint of the residual: g = v^T dR/dQ and v^T dR/d(P0, T0) for the R1..R5 of
// COMPUTE_RESIDUAL (residual.f90:23-62) evaluated after COMPUTE_PRIMITIVE_VARS
// (turbnstwod.f90:204-227), APPLY_BOUNDARY (boundary.f90:24-122, 310-327) and
// COMPUTE_SOURCE_TERMS (source_terms.f90:24-104) on a state Q.
//
// R_i depends on the states of cell i and of its face neighbours only (a ghost
// is built from its parent, which is i, and for type 3 from the neighbour
// across i's opposite face, itself one of i's neighbours).  So dR/dQ is a
// block-sparse matrix with one 5x5 block per (cell, stencil member):
//   k_res_jac  forms, per cell (one thread per dual direction), the five blocks
//              dR_i/dQ_s (s = i, then the neighbours across faces 1..4), and
//              k_res_jac_par dR_i/d(P0, T0), by forward-
//              mode dual numbers through the same physics functions the solver
//              runs (phys.hpp), so the value part is the solver's R bit for bit;
//   k_res_vjp  applies the transpose as a gather: g_k = sum over the (i, s)
//              whose member s is k, in increasing i, of v_i^T dR_i/dQ_s -- the
//              transposed stencil, deterministic, no atomics;
//   k_par_vjp  + k_par_fold: v^T dR/d(P0, T0) as block partial sums folded in a
//              fixed order;
//   k_xy_vjp   v^T dR/d(node x, y): k_res_jac_xy differentiates R_i through the
//              geometry of geometry.f90 recomputed from the nodes it reads
//              (the cell's, its face neighbours', and the ghost centroids'
//              face nodes), and the gather runs over each node's (cell, slot)
//              list.
#include "ctx.hpp"
#include "dual.hpp"
#include "reduce.hpp"

namespace tdt {
namespace {

constexpr int kJacBlocks = 5;         // stencil members per cell: self, faces 1..4
constexpr int kJacStride = 5 * 5 * kJacBlocks;
constexpr int kDualN = 1;             // directions per dual evaluation (per thread)
// node slots per cell: the cell's four nodes, then each face neighbour's four
// (a node shared by several slots collects the sum of their derivatives)
constexpr int kNodeSlots = 20;
constexpr int kXyStride = kNodeSlots * 2 * 5;   // [slot][x, y][m]

// Geometry a cell's residual reads: its surface vectors and volume, and the
// centroids of the cell (index 0) and of the cell or ghost across each face.
template <class R>
struct CellGeo {
    R sc[8];
    R vol;
    R xcc[5], ycc[5];
};

// ... as stored (geometry.f90 arithmetic, computed once at set-up)
template <class R>
__device__ __forceinline__ void geo_stored(const Dev &D, int i, CellGeo<R> &G) {
    const int n = D.n;
    const int4 nb4 = D.nc4[i];
    const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
    for (int k = 0; k < 8; ++k) G.sc[k] = R(D.sc[k * n + i]);
    G.vol = R(D.vol[i]);
    G.xcc[0] = R(D.xcc[i]);
    G.ycc[0] = R(D.ycc[i]);
    for (int f = 0; f < 4; ++f) {
        G.xcc[1 + f] = R(D.xcc[nb[f]]);
        G.ycc[1 + f] = R(D.ycc[nb[f]]);
    }
}

// ... from node coordinates (metres) with k_geometry's and k_geometry_ghosts'
// expressions: x[s], y[s] for the node slots (0-3: the cell's nodes 1-4,
// 4 + 4f + k: node k+1 of the cell across face f+1; unused on ghost faces).
template <class R>
__device__ void geo_from_nodes(const Dev &D, int i, const R (&x)[kNodeSlots],
                               const R (&y)[kNodeSlots], int iaxis, CellGeo<R> &G) {
    const int n = D.n;
    const int4 nb4 = D.nc4[i];
    const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
    const R x1 = x[0], x2 = x[1], x3 = x[2], x4 = x[3];
    const R y1 = y[0], y2 = y[1], y3 = y[2], y4 = y[3];
    G.xcc[0] = (x1 + x2 + x3 + x4) * R(0.25);
    G.ycc[0] = (y1 + y2 + y3 + y4) * R(0.25);
    R vol;
    if (iaxis == 1) {
        G.sc[0] = (y2 - y1) * (y1 + y2) * R(0.5);
        G.sc[1] = -((x2 - x1) * (y1 + y2) * R(0.5));
        G.sc[2] = (y3 - y2) * (y2 + y3) * R(0.5);
        G.sc[3] = -((x3 - x2) * (y2 + y3) * R(0.5));
        G.sc[4] = (y4 - y3) * (y3 + y4) * R(0.5);
        G.sc[5] = -((x4 - x3) * (y3 + y4) * R(0.5));
        G.sc[6] = (y1 - y4) * (y4 + y1) * R(0.5);
        G.sc[7] = -((x1 - x4) * (y4 + y1) * R(0.5));
        vol = R(0.5) * ((x3 - x1) * (y4 - y2) - (x4 - x2) * (y3 - y1)) * G.ycc[0];
    } else {
        G.sc[0] = (y2 - y1);
        G.sc[1] = -(x2 - x1);
        G.sc[2] = (y3 - y2);
        G.sc[3] = -(x3 - x2);
        G.sc[4] = (y4 - y3);
        G.sc[5] = -(x4 - x3);
        G.sc[6] = (y1 - y4);
        G.sc[7] = -(x1 - x4);
        vol = R(0.5) * ((x3 - x1) * (y4 - y2) - (x4 - x2) * (y3 - y1));
    }
    if (vol <= R(0)) vol = -vol;
    G.vol = vol;
    for (int f = 0; f < 4; ++f) {
        if (nb[f] < n) {
            const int b = 4 + 4 * f;
            G.xcc[1 + f] = (x[b] + x[b + 1] + x[b + 2] + x[b + 3]) * R(0.25);
            G.ycc[1 + f] = (y[b] + y[b + 1] + y[b + 2] + y[b + 3]) * R(0.25);
        } else {   // ghost: the centroid mirrored through the face midpoint
            const int a = f, c = (f + 1) & 3;
            const R xf = R(0.5) * (x[a] + x[c]);
            const R yf = R(0.5) * (y[a] + y[c]);
            G.xcc[1 + f] = R(2.0) * xf - G.xcc[0];
            G.ycc[1 + f] = R(2.0) * yf - G.ycc[0];
        }
    }
}

// node ids of the slots of cell i (-1 for the slots of a ghost face)
__device__ __forceinline__ int slot_node(const Dev &D, int i, int s) {
    if (s < 4) return reinterpret_cast<const int *>(D.nod4)[4 * i + s];
    const int f = (s - 4) >> 2, k = (s - 4) & 3;
    const int j = reinterpret_cast<const int *>(D.nc4)[4 * i + f];
    return j < D.n ? reinterpret_cast<const int *>(D.nod4)[4 * j + k] : -1;
}

__device__ __forceinline__ int gid() { return blockIdx.x * kBlock + threadIdx.x; }

// R_i from the states q[0] (cell i) and q[1 + f] (neighbour across face f+1;
// unused for a ghost face), with the same operations in the same order as
// k_primitives, k_bc, k_source and k_residual.
template <class R>
__device__ void cell_residual(const Dev &D, int i, const R (&q)[5][5], const CellGeo<R> &G,
                              const Consts<R> &K, R (&r)[5]) {
    const int n = D.n;
    const int4 nb4 = D.nc4[i];
    const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
    Prim<R> pi, pj[4];
    prim_from_q<R>(pi, q[0][0], q[0][1], q[0][2], q[0][3], q[0][4], K);
    for (int f = 0; f < 4; ++f)
        if (nb[f] < n) prim_from_q<R>(pj[f], q[1 + f][0], q[1 + f][1], q[1 + f][2], q[1 + f][3],
                                      q[1 + f][4], K);
    for (int f = 0; f < 4; ++f) {
        if (nb[f] < n) continue;
        const int k = nb[f] - n, o = (f + 2) & 3;
        const int ty = D.gtype[k];
        R uo = R(0), vo = R(0), to = R(0), po = R(0);
        if (ty == 3) {   // the parent's opposite neighbour: a cell (checked at set-up)
            uo = pj[o].u; vo = pj[o].v; to = pj[o].t; po = pj[o].p;
        }
        ghost_state<R>(pj[f], ty, pi, uo, vo, to, po, G.ycc[1 + f], K);
    }
    // COMPUTE_SOURCE_TERMS
    R s3;
    auto nbr = [&](int f) {
        return src_nbr<R>(G.xcc[1 + f], G.ycc[1 + f], pj[f].u, pj[f].v, pj[f].c[0], pj[f].c[4]);
    };
    const R s5 = source_terms<R>(nbr(0), nbr(1), nbr(2), nbr(3), pi.c[0], pi.c[4], pi.t, pi.p,
                                 pi.v, G.ycc[0], s3, K);
    // COMPUTE_RESIDUAL
    const R(&sc)[8] = G.sc;
    for (int m = 0; m < 5; ++m) {
        R s = (pi.f[m] + pj[0].f[m]) * sc[0] + (pi.g[m] + pj[0].g[m]) * sc[1];
        for (int k = 1; k < 4; ++k) {
            s = s + (pi.f[m] + pj[k].f[m]) * sc[2 * k];
            s = s + (pi.g[m] + pj[k].g[m]) * sc[2 * k + 1];
        }
        r[m] = R(0.5) * s;
    }
    r[4] = r[4] - G.vol * s5;
    if (K.iaxis == 1) r[2] = r[2] - G.vol * s3;
}

// The stencil states of cell i from qs ([5][n], internal order): q[0] = cell,
// q[1 + f] = face neighbour (the cell itself on a ghost face, unused there).
__device__ __forceinline__ void load_stencil(const Dev &D, const double *qs, int i,
                                             double (&q)[5][5]) {
    const int n = D.n;
    const int4 nb4 = D.nc4[i];
    const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
    for (int s = 0; s < 5; ++s) {
        const int j = s == 0 ? i : (nb[s - 1] < n ? nb[s - 1] : i);
        for (int c = 0; c < 5; ++c) q[s][c] = qs[(size_t)c * n + j];
    }
}

// r = R(Q) in FP64 (the forward evaluation behind the adjoint).
__global__ void __launch_bounds__(kBlock) k_res_eval(Dev D, Consts<double> K, const double *qs,
                                                     double *r) {
    const int i = gid();
    if (i >= D.n) return;
    double q[5][5], ri[5];
    load_stencil(D, qs, i, q);
    CellGeo<double> G;
    geo_stored(D, i, G);
    cell_residual<double>(D, i, q, G, K, ri);
    for (int m = 0; m < 5; ++m) r[(size_t)m * D.n + i] = ri[m];
}

// Derivatives of the flow constants with respect to P0 (direction 0) and T0
// (direction 1): S0 = P0^(1-GAMA) (RR T0)^GAMA, T0, TINF = T0 / (1 + (GAMA-1)/2
// AMINF^2), and BL_U = 2 AMINF sqrt(GAMA RR TINF).
struct ParamSeeds {
    double ds0_dp0, ds0_dt0, dtinf_dt0, dblu_dt0;
};

constexpr int kMemberGroups = (5 + kDualN - 1) / kDualN;             // per member
constexpr int kParGroups = (2 + kDualN - 1) / kDualN;                // P0, T0
constexpr int kNodeGroups = (2 * kNodeSlots + kDualN - 1) / kDualN;  // slot x, y

// One dual evaluation per thread, in three kernels (each gets the registers it
// needs alone).  k_res_jac, blockIdx.y = e: member s = e / kMemberGroups (s < 5)
// with components c0.. seeded (J[i] block s = dR_i,m / dQ_s,c [m][c], zero for a
// ghost face).  k_res_jac_par: parameter group e (Jp[i] = [m][p]: dR_i,m /
// d(P0, T0)).  k_res_jac_xy: node group e (Jx[i], directions (slot, x or y)).
__device__ __forceinline__ void res_stencil(const Dev &D, int i, const double *qs,
                                            Dual<kDualN> (&q)[5][5]) {
    double qd[5][5];
    load_stencil(D, qs, i, qd);
    for (int a = 0; a < 5; ++a)
        for (int c = 0; c < 5; ++c) q[a][c] = Dual<kDualN>(qd[a][c]);
}

__global__ void __launch_bounds__(kBlock) k_res_jac(Dev D, Consts<double> K, const double *qs,
                                                    double *J) {
    const int i = gid();
    const int e = blockIdx.y;
    if (i >= D.n) return;
    const int n = D.n;
    const int4 nb4 = D.nc4[i];
    const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
    const int sm = e / kMemberGroups, c0 = (e % kMemberGroups) * kDualN;
    double *Ji = J + (size_t)i * kJacStride + sm * 25;
    if (sm > 0 && nb[sm - 1] >= n) {
        for (int k = 0; k < kDualN && c0 + k < 5; ++k)
            for (int m = 0; m < 5; ++m) Ji[m * 5 + c0 + k] = 0.0;
        return;
    }
    Dual<kDualN> q[5][5], r[5];
    res_stencil(D, i, qs, q);
    const Consts<Dual<kDualN>> KD = lift<kDualN>(K);
    CellGeo<Dual<kDualN>> GD;
    geo_stored(D, i, GD);
    // seeded by static indices (a runtime index would put q in scratch memory)
#pragma unroll
    for (int a = 0; a < 5; ++a)
#pragma unroll
        for (int c = 0; c < 5; ++c)
#pragma unroll
            for (int k = 0; k < kDualN; ++k)
                if (a == sm && c == c0 + k) q[a][c].d[k] = 1.0;
    cell_residual<Dual<kDualN>>(D, i, q, GD, KD, r);
    for (int k = 0; k < kDualN && c0 + k < 5; ++k)
        for (int m = 0; m < 5; ++m) Ji[m * 5 + c0 + k] = r[m].d[k];
}

__global__ void __launch_bounds__(kBlock) k_res_jac_par(Dev D, Consts<double> K, ParamSeeds ps,
                                                        const double *qs, double *Jp) {
    const int i = gid();
    const int e = blockIdx.y;
    if (i >= D.n) return;
    Dual<kDualN> q[5][5], r[5];
    res_stencil(D, i, qs, q);
    // parameter p = e * kDualN + k in direction k
    Consts<Dual<kDualN>> KP = lift<kDualN>(K);
    for (int k = 0; k < kDualN; ++k) {
        const int p = e * kDualN + k;
        if (p == 0) KP.s0.d[k] = ps.ds0_dp0;
        if (p == 1) {
            KP.s0.d[k] = ps.ds0_dt0;
            KP.t0.d[k] = 1.0;
            KP.tinf.d[k] = ps.dtinf_dt0;
            KP.bl_u.d[k] = ps.dblu_dt0;
        }
    }
    CellGeo<Dual<kDualN>> GD;
    geo_stored(D, i, GD);
    cell_residual<Dual<kDualN>>(D, i, q, GD, KP, r);
    for (int k = 0; k < kDualN && e * kDualN + k < 2; ++k)
        for (int m = 0; m < 5; ++m) Jp[(size_t)i * 10 + m * 2 + e * kDualN + k] = r[m].d[k];
}

__global__ void __launch_bounds__(kBlock) k_res_jac_xy(Dev D, Consts<double> K, const double *qs,
                                                       double *Jx) {
    const int i = gid();
    const int e = blockIdx.y;
    if (i >= D.n || e >= kNodeGroups) return;
    Dual<kDualN> q[5][5], r[5];
    res_stencil(D, i, qs, q);
    const Consts<Dual<kDualN>> KD = lift<kDualN>(K);
    const int d0 = e * kDualN;   // node coordinates: this group's directions
    Dual<kDualN> x[kNodeSlots], y[kNodeSlots];
    for (int sl = 0; sl < kNodeSlots; ++sl) {
        const int v = slot_node(D, i, sl);
        x[sl] = Dual<kDualN>(v >= 0 ? D.xn[v] : 0.0);
        y[sl] = Dual<kDualN>(v >= 0 ? D.yn[v] : 0.0);
    }
    // slot (d0 + k) / 2, x if even; seeded by static indices
#pragma unroll
    for (int sl = 0; sl < kNodeSlots; ++sl)
#pragma unroll
        for (int k = 0; k < kDualN; ++k) {
            if (d0 + k == 2 * sl) x[sl].d[k] = 1.0;
            if (d0 + k == 2 * sl + 1) y[sl].d[k] = 1.0;
        }
    CellGeo<Dual<kDualN>> GX;
    geo_from_nodes(D, i, x, y, K.iaxis, GX);
    cell_residual<Dual<kDualN>>(D, i, q, GX, KD, r);
    double *Jxi = Jx + (size_t)i * kXyStride;
    for (int k = 0; k < kDualN && d0 + k < 2 * kNodeSlots; ++k)
        for (int m = 0; m < 5; ++m) Jxi[(d0 + k) * 5 + m] = r[m].d[k];
}

// gxy_v = sum over the (cell i, slot s) whose slot holds node v, in order, of
// v_i^T dR_i/d(x, y)_s
__global__ void __launch_bounds__(kBlock) k_xy_vjp(Dev D, const double *Jx, const double *v,
                                                   const int *nptr, const int2 *nlist,
                                                   double *gxy) {
    const int nd = gid();
    if (nd >= D.m) return;
    const int n = D.n;
    double ax = 0.0, ay = 0.0;
    for (int e = nptr[nd]; e < nptr[nd + 1]; ++e) {
        const int2 is = nlist[e];
        const double *B = Jx + (size_t)is.x * kXyStride + is.y * 10;
        double tx = v[is.x] * B[0], ty = v[is.x] * B[5];
        for (int m = 1; m < 5; ++m) {
            const double vm = v[(size_t)m * n + is.x];
            tx = tx + vm * B[m];
            ty = ty + vm * B[5 + m];
        }
        ax = ax + tx;
        ay = ay + ty;
    }
    gxy[nd] = ax;
    gxy[D.m + nd] = ay;
}

// g_k = sum over the entries (i, s) of cell k's transposed stencil, in order,
// of v_i^T dR_i/dQ_s.
__global__ void __launch_bounds__(kBlock) k_res_vjp(Dev D, const double *J, const double *v,
                                                    const int *tptr, const int2 *tlist,
                                                    double *g) {
    const int k = gid();
    if (k >= D.n) return;
    const int n = D.n;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int e = tptr[k]; e < tptr[k + 1]; ++e) {
        const int2 is = tlist[e];
        const double *B = J + (size_t)is.x * kJacStride + is.y * 25;
        double vi[5];
        for (int m = 0; m < 5; ++m) vi[m] = v[(size_t)m * n + is.x];
        for (int c = 0; c < 5; ++c) {
            double t = vi[0] * B[c];
            for (int m = 1; m < 5; ++m) t = t + vi[m] * B[m * 5 + c];
            acc[c] = acc[c] + t;
        }
    }
    for (int c = 0; c < 5; ++c) g[(size_t)c * n + k] = acc[c];
}

// per block: the sums over its cells of v_i^T dR_i/d(P0, T0)
__global__ void __launch_bounds__(kBlock) k_par_vjp(Dev D, const double *Jp, const double *v,
                                                    double *part) {
    const int i = gid();
    double t[2] = {0.0, 0.0};
    if (i < D.n)
        for (int p = 0; p < 2; ++p) {
            double a = v[i] * Jp[(size_t)i * 10 + p];
            for (int m = 1; m < 5; ++m) a = a + v[(size_t)m * D.n + i] * Jp[(size_t)i * 10 + m * 2 + p];
            t[p] = a;
        }
    __shared__ double sh[kBlock / kWave][2];
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    for (int p = 0; p < 2; ++p) {
        const double s = wave_sum(t[p]);
        if (lane == 0) sh[w][p] = s;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double a = sh[0][threadIdx.x];
        for (int q = 1; q < kBlock / kWave; ++q) a = a + sh[q][threadIdx.x];
        part[blockIdx.x * 2 + threadIdx.x] = a;
    }
}

// one thread per parameter: the block partials in block order
__global__ void k_par_fold(const double *part, int nblk, double *out) {
    const int p = threadIdx.x;
    if (p >= 2) return;
    double a = 0.0;
    for (int b = 0; b < nblk; ++b) a = a + part[b * 2 + p];
    out[p] = a;
}

}  // namespace

hipError_t launch_res_eval(tdt_ctx *c, const double *qs, double *r) {
    hipLaunchKernelGGL(k_res_eval, dim3(nblocks(c->D.n)), dim3(kBlock), 0, c->stream, c->D, c->K,
                       qs, r);
    return hipGetLastError();
}

static ParamSeeds param_seeds(const tdt_ctx *c) {
    const Consts<double> &K = c->K;
    const tdt_params &P = c->params;
    ParamSeeds ps;
    const double gam = P.gama;
    ps.ds0_dp0 = (1.0 - gam) * (double)K.s0 / P.p0;
    ps.ds0_dt0 = gam * (double)K.s0 / P.t0;
    ps.dtinf_dt0 = (double)K.tinf / P.t0;
    ps.dblu_dt0 = 0.5 * (double)K.bl_u / P.t0;
    return ps;
}

hipError_t launch_res_jac(tdt_ctx *c, const double *qs, double *J, double *Jp, double *Jx) {
    const unsigned nb = nblocks(c->D.n);
    hipLaunchKernelGGL(k_res_jac, dim3(nb, kJacBlocks * kMemberGroups), dim3(kBlock), 0,
                       c->stream, c->D, c->K, qs, J);
    hipLaunchKernelGGL(k_res_jac_par, dim3(nb, kParGroups), dim3(kBlock), 0, c->stream, c->D,
                       c->K, param_seeds(c), qs, Jp);
    if (Jx)
        hipLaunchKernelGGL(k_res_jac_xy, dim3(nb, kNodeGroups), dim3(kBlock), 0, c->stream, c->D,
                           c->K, qs, Jx);
    return hipGetLastError();
}

hipError_t launch_res_apply(tdt_ctx *c, const double *J, const double *Jp, const double *v,
                            const int *tptr, const int2 *tlist, double *part, double *g,
                            double *gp) {
    const unsigned nb = nblocks(c->D.n);
    hipLaunchKernelGGL(k_res_vjp, dim3(nb), dim3(kBlock), 0, c->stream, c->D, J, v, tptr, tlist,
                       g);
    if (gp) {   // gp == nullptr: the state gradient only
        hipLaunchKernelGGL(k_par_vjp, dim3(nb), dim3(kBlock), 0, c->stream, c->D, Jp, v, part);
        hipLaunchKernelGGL(k_par_fold, dim3(1), dim3(kWave), 0, c->stream, part, (int)nb, gp);
    }
    return hipGetLastError();
}

hipError_t launch_res_vjp(tdt_ctx *c, const double *qs, const double *v, const int *tptr,
                          const int2 *tlist, double *J, double *Jp, double *part, double *g,
                          double *gp, double *Jx, const int *nptr, const int2 *nlist,
                          double *gxy) {
    hipError_t e = launch_res_jac(c, qs, J, Jp, gxy ? Jx : nullptr);
    if (e == hipSuccess) e = launch_res_apply(c, J, Jp, v, tptr, tlist, part, g, gp);
    if (e == hipSuccess && gxy)
        hipLaunchKernelGGL(k_xy_vjp, dim3(nblocks(c->D.m)), dim3(kBlock), 0, c->stream, c->D, Jx,
                           v, nptr, nlist, gxy);
    return e == hipSuccess ? hipGetLastError() : e;
}

}  // namespace tdt
