// Device code of the fused hot path, included by fused.hip and fused_mixed.hip
// (each instantiates one precision).  See fused.hip for the kernel sequence.
#pragma once

#include "fused.hpp"
#include "quad.hpp"

namespace tdt {
namespace {

// Logical block within a range of nb8 blocks (nb8 and the range's first block
// multiples of 8).  Blocks are dispatched round-robin over the 8 XCDs, so this
// gives each XCD a contiguous eighth of the range: its L2 then holds most of
// the neighbours and nodes its cells gather.
__device__ __forceinline__ int xcd_chunk(int local, int nb8) {
    return (local & 7) * (nb8 >> 3) + (local >> 3);
}

// Logical block for a grid of any size (the quad kernels): XCD x gets blocks
// [x * per + min(x, rem), ...) with per = nb / 8, rem = nb % 8.
__device__ __forceinline__ int logical_block(int remap) {
    int b = blockIdx.x, nb = gridDim.x;
    if (!remap || nb < 16) return b;
    int xcd = b & 7, local = b >> 3, per = nb >> 3, rem = nb & 7;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// Logical cell block of a cell kernel launched on F.ncb8 blocks (the blocks
// past the last cell block have no cells).
__device__ __forceinline__ int cell_block(const FDev &F) {
    return F.remap ? xcd_chunk(blockIdx.x, F.ncb8) : blockIdx.x;
}

// FP32-mixed stores the frozen dissipation and viscous terms combined, E1 = D1,
// Eq = Dq + VIS(q-1) (the stage updates subtract them together), five values
// per cell instead of nine; FP64 keeps D and VIS apart, as the reference
// subtracts them one after the other.
template <class SD>
constexpr bool kCombinedDV = sizeof(SD) == 4;

template <class T>
__device__ __forceinline__ double ld(const void *p, size_t i) {
    return (double)static_cast<const T *>(p)[i];
}
template <class T>
__device__ __forceinline__ void stv(void *p, size_t i, double v) {
    static_cast<T *>(p)[i] = (T)v;
}

// Records (drec, nrec): kRecLen values of type T per cell or node, stored as
// planes of 16-byte vectors: plane k holds values 4k..4k+3 (FP32) or 2k, 2k+1
// (FP64) of every record, plane stride = the record count.  A wave's record
// store of one plane is one contiguous 1 KB run, and a gather of a record
// costs two (FP32) or four (FP64) 16-byte loads.
template <class T>
__device__ __forceinline__ void rec_get(const void *p, size_t stride, size_t i,
                                        T (&r)[kRecLen]) {
    if constexpr (sizeof(T) == 4) {
        const float4 *q = static_cast<const float4 *>(p);
        const float4 a = q[i], b = q[stride + i];
        r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w;
        r[4] = b.x; r[5] = b.y; r[6] = b.z; r[7] = b.w;
    } else {
        const double2 *q = static_cast<const double2 *>(p);
        const double2 a = q[i], b = q[stride + i], c = q[2 * stride + i], d = q[3 * stride + i];
        r[0] = a.x; r[1] = a.y; r[2] = b.x; r[3] = b.y;
        r[4] = c.x; r[5] = c.y; r[6] = d.x; r[7] = d.y;
    }
}
template <class T>
__device__ __forceinline__ void rec_put(void *p, size_t stride, size_t i, const T (&r)[kRecLen]) {
    if constexpr (sizeof(T) == 4) {
        float4 *q = static_cast<float4 *>(p);
        q[i] = make_float4(r[0], r[1], r[2], r[3]);
        q[stride + i] = make_float4(r[4], r[5], r[6], r[7]);
    } else {
        double2 *q = static_cast<double2 *>(p);
        q[i] = make_double2(r[0], r[1]);
        q[stride + i] = make_double2(r[2], r[3]);
        q[2 * stride + i] = make_double2(r[4], r[5]);
        q[3 * stride + i] = make_double2(r[6], r[7]);
    }
}
// Value `slot` of record i (single-value stores of the quad pre-pass).
template <class T>
__device__ __forceinline__ T &rec_slot(void *p, size_t stride, size_t i, int slot) {
    constexpr int w = 16 / sizeof(T);   // values per 16-byte vector
    return static_cast<T *>(p)[((size_t)(slot / w) * stride + i) * w + slot % w];
}

// Node averages (UN, VN, TN, RHON, RN) of node v: slots 0..3 in 16-byte
// vector planes, slot 4 in a plain array after them.
template <class T>
__device__ __forceinline__ T &nav_slot(void *p, size_t m, size_t v, int slot) {
    if (slot == 4) return static_cast<T *>(p)[4 * m + v];
    constexpr int w = 16 / sizeof(T);
    return static_cast<T *>(p)[((size_t)(slot / w) * m + v) * w + slot % w];
}
template <class T>
__device__ __forceinline__ void nav_put(void *p, size_t m, size_t v, const T (&a)[5]) {
    if constexpr (sizeof(T) == 4) {
        static_cast<float4 *>(p)[v] = make_float4(a[0], a[1], a[2], a[3]);
    } else {
        static_cast<double2 *>(p)[v] = make_double2(a[0], a[1]);
        static_cast<double2 *>(p)[m + v] = make_double2(a[2], a[3]);
    }
    static_cast<T *>(p)[4 * m + v] = a[4];
}
// ... and with the node's coordinates
template <class T>
__device__ __forceinline__ void nav_get(const FDev &F, size_t v, T (&a)[5], T &x, T &y) {
    const size_t m = F.m;
    if constexpr (sizeof(T) == 4) {
        const float4 q = static_cast<const float4 *>(F.nrec)[v];
        a[0] = q.x; a[1] = q.y; a[2] = q.z; a[3] = q.w;
        const float2 c = static_cast<const float2 *>(F.nxy)[v];
        x = c.x; y = c.y;
    } else {
        const double2 q0 = static_cast<const double2 *>(F.nrec)[v];
        const double2 q1 = static_cast<const double2 *>(F.nrec)[m + v];
        a[0] = q0.x; a[1] = q0.y; a[2] = q1.x; a[3] = q1.y;
        const double2 c = static_cast<const double2 *>(F.nxy)[v];
        x = c.x; y = c.y;
    }
    a[4] = static_cast<const T *>(F.nrec)[4 * m + v];
}

// VOL/DT of a cell from its record slot: the pre-pass stores VOL/DT when each
// cell has its own time step, VOL when all share DTMIN (known only after the
// pre-pass), here dtmin.
template <class R, class SD>
__device__ __forceinline__ R rec_vbydt(const SD (&r)[kRecLen], int local, double dtmin) {
    return local ? (R)r[kRecVbydt] : fdiv((R)r[kRecVbydt], (R)dtmin);
}

// The dissipation D of cell i (solver_state.f90:183-223) from the records of i
// and its four face neighbours (a ghost neighbour is a copy of the cell) and
// the FP64 states: returns D1..D5 in d (2nd-difference minus 4th-difference
// sums, each accumulated in face order).  KEEP: also returns the neighbours'
// states, converted to R, in qk (a ghost's is the cell's own), for a caller
// that needs them next (stage 1's flux faces read the same Q^n).
template <class R, class SD, bool KEEP = false>
__device__ __forceinline__ void dissipation_rec(const FDev &F, int i, const int (&nb)[4],
                                                const double (&ci)[5], int local, double dti,
                                                R (&d)[5], R (*qk)[5] = nullptr) {
    const int nt = F.nt, n = F.n;
    const double *Q = F.qn;
    SD ri[kRecLen];
    rec_get<SD>(F.drec, F.n, i, ri);
    const R volbydt = rec_vbydt<R>(ri, local, dti);
    R d2[5], d4[5];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const int j = nb[f] < n ? nb[f] : i;
        SD rj[kRecLen];
        rec_get<SD>(F.drec, F.n, j, rj);
        const R vbn = rec_vbydt<R>(rj, local, dti);
        R a2, a4, b2, b4;
        jst_coeffs_v<R>((R)ri[kRecAnu], (R)rj[kRecAnu], volbydt, vbn, a2, a4);
        jst_coeffs_v<R>((R)ri[kRecAnurt], (R)rj[kRecAnurt], volbydt, vbn, b2, b4);
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const double qj = Q[q * nt + j];
            if constexpr (KEEP) qk[f][q] = (R)qj;
            const R t2 = (q < 4 ? a2 : b2) * (R)(qj - ci[q]);
            const R t4 = (q < 4 ? a4 : b4) * ((R)rj[kRecD2q + q] - (R)ri[kRecD2q + q]);
            d2[q] = f == 0 ? t2 : d2[q] + t2;
            d4[q] = f == 0 ? t4 : d4[q] + t4;
        }
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) d[q] = d2[q] - d4[q];
}

// Surface vector of face f of cell i, from node bn to node dn.  FP64 forms it
// from the two nodes' coordinates with COMPUTE_GEOMETRY's operations
// (geometry.f90:77-97), which gives the SC arrays' bits without streaming them,
// while those arrays are the nodes' (F.sc_nodes: 1 planar, 2 axisymmetric);
// after tdt_set_array has written SC, and in FP32-mixed (whose FP32 node
// coordinates would lose the short wall faces' length), it reads SC.  NODES:
// the caller gathers the face's nodes anyway (k_dv).  (The stage kernels read
// SC: their flux faces would wait on the node ids before the node gathers,
// which costs more on latency-bound grids than the streamed SC.)
template <class R, class SD, bool NODES>
__device__ __forceinline__ void face_sc(const FDev &F, int f, int i, int bn, int dn, R &sx, R &sy) {
    if constexpr (NODES && sizeof(SD) == 8) {
        if (F.sc_nodes) {
            const double2 b = static_cast<const double2 *>(F.nxy)[bn];
            const double2 d = static_cast<const double2 *>(F.nxy)[dn];
            if (F.sc_nodes == 2) {
                sx = (d.y - b.y) * (b.y + d.y) * 0.5;
                sy = -((d.x - b.x) * (b.y + d.y) * 0.5);
            } else {
                sx = d.y - b.y;
                sy = -(d.x - b.x);
            }
            return;
        }
    }
    sx = (R)ld<SD>(F.gsc, (size_t)(2 * f) * F.n + i);
    sy = (R)ld<SD>(F.gsc, (size_t)(2 * f + 1) * F.n + i);
}

// The viscous terms of face f of cell A (viscous.f90:130-217): B, D = the
// face's nodes (their records: averages and coordinates), C = the neighbour.
template <class R, class SD>
__device__ __forceinline__ ViscFace<R> viscous_face_rec(const FDev &F, R sx, R sy, R xa, R ya,
                                                        R ua, R va, R ta, R rhoa, R ra, int bn,
                                                        int dn, R xc, R yc, R uc, R vc, R tc,
                                                        R rhoc, R rc, const Consts<R> &KR) {
    SD b[5], e[5], xb, yb, xd, yd;
    nav_get<SD>(F, bn, b, xb, yb);
    nav_get<SD>(F, dn, e, xd, yd);
    return viscous_face<R>(sx, sy, xa, ya, (R)xb, (R)yb, xc, yc, (R)xd, (R)yd, ua, va, ta, rhoa,
                           ra, (R)b[0], (R)b[1], (R)b[2], (R)b[3], (R)b[4], uc, vc, tc, rhoc, rc,
                           (R)e[0], (R)e[1], (R)e[2], (R)e[3], (R)e[4], KR);
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

// Batch: folds case cs's pending partials (one block per case) into
// hist[it][cs].  The iteration counter moves on in the next stage 4 (or
// k_batch_tick), after every case has read it.
__device__ __forceinline__ void fold_case(const FDev &F, int cs) {
    if (F.ctr[3] == 0) return;   // uniform across the block
    __shared__ int flag;
    if (threadIdx.x == 0) flag = 0;
    __syncthreads();
    const int it = F.ctr[0];
    const int4 r = F.ccase[cs];
    reduce_partials(F.part + (size_t)r.x * kPartStride, r.y, r.z,
                    F.hist + ((size_t)it * F.ncases + cs) * 10, &flag);
    __syncthreads();
    if (threadIdx.x == 0 && flag) atomicCAS(&F.ctr[2], 0, it + 1);
}

// Batch: the flow constants of case cs in place of the kernel arguments'.  The
// cases of a batch share the gas and model constants (checked when the batch is
// made), so only the fields of the case's state and boundary data are loaded.
template <class R>
__device__ __forceinline__ void case_fields(Consts<R> &K, const Consts<R> &C) {
    K.t0 = C.t0; K.s0 = C.s0; K.pamb = C.pamb; K.pamb_set = C.pamb_set;
    K.pinf = C.pinf; K.tinf = C.tinf; K.bl_u = C.bl_u; K.deltabl = C.deltabl;
    K.rtol = C.rtol; K.rtol_rtred = C.rtol_rtred;
}

template <bool BT, class R>
__device__ __forceinline__ void case_consts(const FDev &F, int cs, Consts<double> &K,
                                            Consts<R> &KR) {
    if constexpr (BT) {
        case_fields(K, static_cast<const Consts<double> *>(F.kc64)[cs]);
        case_fields(KR, static_cast<const Consts<R> *>(F.kcr)[cs]);
    }
}

// BC type 6 ghost k (parent par) from the cells' states in q (storage type T,
// component stride qs): the port column's C1, C2, the parent face's surface
// vector and the centroids, all as the flux arithmetic reads them.
template <class R, class T, class SD = R>
__device__ __forceinline__ void bc6_fused(const FDev &F, Prim<R> &s, const Prim<R> &pp, int k,
                                          int par, const void *q, size_t qs,
                                          const Consts<R> &KR) {
    const int f = F.gside[k] - 1, n = F.n;
    const int gid = n + k;
    bc6_state<R>(s, pp, (R)ld<SD>(F.gsc, (size_t)(2 * f) * n + par),
                 (R)ld<SD>(F.gsc, (size_t)(2 * f + 1) * n + par),
                 (R)fabs(ld<SD>(F.gycc, gid) + ld<SD>(F.gycc, par)), F.g6ptr[k], F.g6ptr[k + 1],
                 F.g6cell, F.g6dely,
                 [&](int c, R &c1, R &c2, R &yc) {
                     c1 = (R)ld<T>(q, c);
                     c2 = (R)ld<T>(q, qs + c);
                     yc = (R)ld<SD>(F.gycc, c);
                 },
                 KR);
}

// k_bc6: the type-6 ghosts of one stage's input state, each from its parent and
// its port column (a walk across the grid), formed once per stage so that the
// cell kernels only load them: F.g6s[6][g] = U, V, T, P, rho, C5 in the
// arithmetic type R (other ghosts' entries are not written).
template <class R, class SI, class SD, int STAGE>
__global__ void __launch_bounds__(kBlock) k_bc6(FDev F, Consts<R> KR) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= F.g || F.gtype[k] != 6) return;
    const void *qin = STAGE == 1 ? (const void *)F.qn : (const void *)F.st[STAGE & 1];
    const size_t sin = STAGE == 1 ? (size_t)F.nt : (size_t)F.n;
    const int par = F.gparent[k], g = F.g;
    Prim<R> pp, s;
    prim_from_q<R>(pp, (R)ld<SI>(qin, par), (R)ld<SI>(qin, sin + par),
                   (R)ld<SI>(qin, 2 * sin + par), (R)ld<SI>(qin, 3 * sin + par),
                   (R)ld<SI>(qin, 4 * sin + par), KR);
    bc6_fused<R, SI, SD>(F, s, pp, k, par, qin, sin, KR);
    stv<R>(F.g6s, k, s.u);
    stv<R>(F.g6s, (size_t)g + k, s.v);
    stv<R>(F.g6s, (size_t)2 * g + k, s.t);
    stv<R>(F.g6s, (size_t)3 * g + k, s.p);
    stv<R>(F.g6s, (size_t)4 * g + k, s.c[0]);
    stv<R>(F.g6s, (size_t)5 * g + k, s.c[4]);
}

// The type-6 ghost k as k_bc6 left it: the values a node gather reads ...
template <class R>
__device__ __forceinline__ void bc6_load(const FDev &F, int k, R &u, R &v, R &t, R &rho, R &c5) {
    const size_t g = F.g;
    u = (R)ld<R>(F.g6s, k);
    v = (R)ld<R>(F.g6s, g + k);
    t = (R)ld<R>(F.g6s, 2 * g + k);
    rho = (R)ld<R>(F.g6s, 4 * g + k);
    c5 = (R)ld<R>(F.g6s, 5 * g + k);
}

// ... and all six, for ghost_state.
template <class R>
struct Bc6Load {
    const FDev &F;
    int k;
    __device__ void operator()(R &u, R &v, R &t, R &p, R &rho, R &c5) const {
        bc6_load<R>(F, k, u, v, t, rho, c5);
        p = (R)ld<R>(F.g6s, (size_t)3 * F.g + k);
    }
};

// The loader ghost_state takes for ghost k: the type-6 states when the grid has
// them (B6); otherwise none, and the kernels compile as for a grid without.
template <bool B6, class R>
__device__ __forceinline__ auto g6_for(const FDev &F, int k) {
    if constexpr (B6)
        return Bc6Load<R>{F, k};
    else
        return NoBc6{};
}

// Node gather through the contributor list, for nodes with more than
// kNodeWidth contributors (viscous.f90:77-124 in the reference's order).
template <class R, bool B6>
__device__ __forceinline__ void node_gather_list(const FDev &F, const Consts<R> &KR, int v, R &au,
                                                 R &av, R &at, R &arho, R &ar) {
    const int nt = F.nt;
    const double *Q = F.qn;
    int4 ov = F.novr[v];
    int j0 = F.nptr[v], j1 = F.nptr[v + 1];
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
            if (B6 && ty == 6) {
                bc6_load<R>(F, k, u, vv, t, rho, c5);
            } else {
                ghost_state<R>(s, ty, pp, uo, vo, to, po, (R)F.ycc[id], KR);
                u = s.u; vv = s.v; t = s.t; rho = s.c[0]; c5 = s.c[4];
            }
        }
        R rr = fdiv(c5, rho);
        if (j >= ov.x) au += u;
        if (j >= ov.y) av += vv;
        at += t;
        arho += rho;
        if (j >= ov.z) ar += rr;
    }
}

// ---------------------------------------------------------------------------------
// k_pre: ncb8 cell blocks and nnb8 node blocks (both padded to multiples of 8
// and spread so that each XCD works on contiguous cells and the nodes around
// them, alternating between the two), after a first block of its own that folds
// the previous iteration's monitor partials (off every cell's path; one per case
// in a batch).  Cells:
// primitives of Q^n, DT, and the dissipation record (sensors, VOL/DT, D2Q);
// nodes: the nodal averages.
template <class SD, class R, bool B6, bool BT>
__global__ void __launch_bounds__(kBlock) k_pre(FDev F, Consts<double> K, Consts<R> KR,
                                                int nb_partials, int ncb8, int nnb8) {
    // The folding blocks (one per case of a batch) come first: dispatched with the
    // first blocks, they run beside the cell and node blocks, where as the last
    // blocks their fold outlasted them (FP64 ramp: 73 -> 64 us).  Their count is
    // padded to a multiple of 8, which keeps every other block on the XCD that the
    // mapping below assumes (blocks are dispatched round-robin over the 8 XCDs).
    int b = blockIdx.x;
    {
        const int nfold = BT ? F.ncases : 1, nfold8 = (nfold + 7) & ~7;
        if (b < nfold8) {
            if (b < nfold) {
                if constexpr (BT)
                    fold_case(F, b);
                else
                    fold_monitor(F, nb_partials);
            }
            return;
        }
        b -= nfold8;
    }
    // Each XCD takes its eighth of the cell blocks and of the node blocks,
    // alternating between them, so that a node block runs while the states of
    // the cells around its nodes are still in that XCD's L2.
    bool cells;
    int lcb;
    if (F.remap) {
        const int xcd = b & 7, l = b >> 3, cx = ncb8 >> 3, nx = nnb8 >> 3;
        const int both = 2 * min(cx, nx);
        if (l < both) {
            cells = (l & 1) == 0;
            lcb = xcd * (cells ? cx : nx) + (l >> 1);
        } else {
            cells = cx > nx;
            lcb = xcd * (cells ? cx : nx) + (l - both + both / 2);
        }
    } else {
        cells = b < ncb8;
        lcb = cells ? b : b - ncb8;
    }
    if (cells) {
        const int cb = lcb;
        int i;
        bool act;
        if constexpr (BT) {
            const int4 bl = F.cblk[cb];
            if (bl.y == 0) return;
            case_consts<BT>(F, bl.z, K, KR);
            i = bl.x + threadIdx.x;
            act = (int)threadIdx.x < bl.y;
        } else {
            if (cb * kBlock >= F.n) return;
            i = cb * kBlock + threadIdx.x;
            act = i < F.n;
        }
        double dtl = 100000.0;
        if (act) {
            const int nt = F.nt;
            const double *Q = F.qn;
            double c0 = Q[i], c1 = Q[nt + i], c2 = Q[2 * nt + i], c3 = Q[3 * nt + i],
                   c4 = Q[4 * nt + i];
            double u, v, t, p;
            uvtp_from_q(c0, c1, c2, c3, u, v, t, p, K);
            double vel = sqrt(u * u + v * v);
            double dt = fdiv(K.cfl * F.dl[i], fabs(vel) + sqrt(K.gamma_rr * t));
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
            SD rec[kRecLen];
            rec[kRecAnu] = (SD)fdiv(anun, anud);
            anun = fabs(cn[4][0] - c4) + fabs(cn[4][1] - c4) + fabs(cn[4][2] - c4) +
                   fabs(cn[4][3] - c4);
            anud = fabs(cn[4][0] + c4) + fabs(cn[4][1] + c4) + fabs(cn[4][2] + c4) +
                   fabs(cn[4][3] + c4);
            rec[kRecAnurt] = (SD)fdiv(anun, dmax(anud, 8.0 * c0 * K.rtol));
            // VOL/DT as the dissipation faces divide them (R arithmetic), or VOL
            // when the time step is the global minimum
            const R vol = (R)ld<SD>(F.gvol, i);
            rec[kRecVbydt] = (SD)(K.local ? fdiv(vol, (R)dt) : vol);
            const double ci[5] = {c0, c1, c2, c3, c4};
            for (int m = 0; m < 5; ++m)
                rec[kRecD2q + m] = (SD)(cn[m][0] + cn[m][1] + cn[m][2] + cn[m][3] - 4.0 * ci[m]);
            rec_put<SD>(F.drec, F.n, i, rec);
        }
        if (!K.local) block_atomic_min_pos(dtl, reinterpret_cast<double *>(F.dtslot));
        return;
    }
    const int nbl = lcb;
    int v;
    if constexpr (BT) {
        const int4 bl = F.nblk[nbl];
        if ((int)threadIdx.x >= bl.y) return;
        case_consts<BT>(F, bl.z, K, KR);
        v = bl.x + threadIdx.x;
    } else {
        v = nbl * kBlock + threadIdx.x;
        if (v >= F.m) return;
    }
    const int hr = F.nrel[v];
    const int cnt = hr & 15;
    R au = R(0), av = R(0), at = R(0), arho = R(0), ar = R(0);
    if (cnt <= kNodeWidth) {
        // fixed-width records: the contributing cells' states are requested
        // before any is used (a ghost's parent follows in its own branch)
        const int nt = F.nt, n = F.n;
        const double *Q = F.qn;
        const int su = (hr >> 4) & 15, sv = (hr >> 8) & 15, sr = (hr >> 12) & 15;
        ar = (hr >> 16) ? KR.rtol_rtred : R(0);
        int2 e[kNodeWidth];
        R q[kNodeWidth][5];
#pragma unroll
        for (int k = 0; k < kNodeWidth; ++k) e[k] = F.ncon[(size_t)v * kNodeWidth + k];
#pragma unroll
        for (int k = 0; k < kNodeWidth; ++k)
#pragma unroll
            for (int c = 0; c < 5; ++c) q[k][c] = (R)Q[(size_t)c * nt + e[k].y];
#pragma unroll
        for (int k = 0; k < kNodeWidth; ++k) {
            if (k >= cnt) break;
            R u, vv, t, rho, c5;
            if (e[k].x < n) {
                R p;
                rho = q[k][0];
                uvtp_from_q<R>(rho, q[k][1], q[k][2], q[k][3], u, vv, t, p, KR);
                c5 = q[k][4];
            } else {
                // a ghost from its parent's state (loaded above as the source cell)
                const int a = e[k].x & kIdMask, ty = (unsigned)e[k].x >> kTypeShift;
                Prim<R> pp, s;
                prim_from_q<R>(pp, q[k][0], q[k][1], q[k][2], q[k][3], q[k][4], KR);
                R uo = R(0), vo = R(0), to = R(0), po = R(0), ycg = R(0);
                if (ty == 3)
                    uvtp_from_q<R>((R)Q[a], (R)Q[nt + a], (R)Q[2 * nt + a], (R)Q[3 * nt + a], uo,
                                   vo, to, po, KR);
                else
                    ycg = (R)ld<SD>(F.gycc, a);
                if (B6 && ty == 6) {
                    bc6_load<R>(F, a - n, u, vv, t, rho, c5);
                } else {
                    ghost_state<R>(s, ty, pp, uo, vo, to, po, ycg, KR);
                    u = s.u; vv = s.v; t = s.t; rho = s.c[0]; c5 = s.c[4];
                }
            }
            const R rr = fdiv(c5, rho);
            if (k >= su) au += u;
            if (k >= sv) av += vv;
            at += t;
            arho += rho;
            if (k >= sr) ar += rr;
        }
    } else {
        ar = F.novr[v].w ? KR.rtol_rtred : R(0);
        node_gather_list<R, B6>(F, KR, v, au, av, at, arho, ar);
    }
    const R nc = (R)(cnt <= kNodeWidth ? cnt : F.nptr[v + 1] - F.nptr[v]);
    const Recip<R> rcnt(nc);
    const SD a[5] = {(SD)rcnt.div(au), (SD)rcnt.div(av), (SD)rcnt.div(at), (SD)rcnt.div(arho),
                     (SD)dmax(KR.rtol, rcnt.div(ar))};
    nav_put<SD>(F.nrec, F.m, v, a);
}

// ---------------------------------------------------------------------------------
// Node records of a wavefront's cells staged in LDS (FP64 k_dv with the node
// tables): per wave, NR rounds of 64 slots of UN, VN | TN, RHON | X, Y as 16-byte
// vectors and RN as two 32-bit halves, 56 bytes per slot, in dynamic LDS (the
// block's four waves' regions one after the other).
struct NodeLds {
    double2 *uv, *tr, *xy;
    unsigned *rlo, *rhi;
};

constexpr size_t kNodeLdsBytes = 3 * sizeof(double2) + 2 * sizeof(unsigned);   // per slot

__device__ __forceinline__ NodeLds node_lds(int slots) {
    extern __shared__ __align__(16) unsigned char node_smem[];
    NodeLds s;
    s.uv = reinterpret_cast<double2 *>(node_smem);
    s.tr = s.uv + slots;
    s.xy = s.tr + slots;
    s.rlo = reinterpret_cast<unsigned *>(s.xy + slots);
    s.rhi = s.rlo + slots;
    return s;
}

// Direct global-to-LDS loads: lane l of the wave writes its 16 (4) bytes at
// dst_wave + l, so dst_wave must be the same for the whole wave.
template <class T>
__device__ __forceinline__ void lds_copy16(const T *src, T *dst_wave) {
    __builtin_amdgcn_global_load_lds(static_cast<const void *>(src),
                                     (__attribute__((address_space(3))) void *)dst_wave, 16, 0, 0);
}

__device__ __forceinline__ void lds_copy4(const unsigned *src, unsigned *dst_wave) {
    __builtin_amdgcn_global_load_lds(static_cast<const void *>(src),
                                     (__attribute__((address_space(3))) void *)dst_wave, 4, 0, 0);
}

// The node ids of this wave's NR copy rounds: slot k = r * 64 + lane in round r
// (F.nl_cap = NR * 64 slots per wave, the list padded to whole rounds).
template <int NR>
__device__ __forceinline__ void node_lds_ids(const FDev &F, int lb, int (&v)[NR]) {
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int *bn = F.bnode + ((size_t)lb * (kBlock / kWave) + w) * (NR * kWave) + lane;
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] = bn[r * kWave];
}

// Waits for the ids (and whatever else is in flight: one wait for all, since a
// wait between copies would also wait for the copies issued before it), then
// issues this wave's copies; a round's 64 lanes write 64 consecutive slots.
template <int NR>
__device__ __forceinline__ void node_lds_copy(const FDev &F, const NodeLds &s, const int (&v)[NR]) {
    const size_t m = F.m;
    const double2 *uv = static_cast<const double2 *>(F.nrec);
    const double2 *tr = uv + m;
    const unsigned *rn =
        reinterpret_cast<const unsigned *>(static_cast<const double *>(F.nrec) + 4 * m);
    const double2 *xy = static_cast<const double2 *>(F.nxy);
    const int w = threadIdx.x / kWave;
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0); expcnt, lgkmcnt unconstrained
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int ws = (w * NR + r) * kWave;
        lds_copy16(uv + v[r], s.uv + ws);
        lds_copy16(tr + v[r], s.tr + ws);
        lds_copy16(xy + v[r], s.xy + ws);
        lds_copy4(rn + 2 * (size_t)v[r], s.rlo + ws);
        lds_copy4(rn + 2 * (size_t)v[r] + 1, s.rhi + ws);
    }
}

// This wave's copies are in LDS once its vector memory count is zero; a wave
// reads only its own region, so no block barrier is needed.
__device__ __forceinline__ void node_lds_wait() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// k_dv: dissipation D and viscous fluxes VIS from Q^n in a kernel of their own
// (the split variant; the merged variant forms them inside stage 1).  NR > 0
// (k_dv_nodes: FP64 single case, LOCAL = 1): the viscous faces read the nodes
// from the wave's LDS table, copied in NR rounds while the dissipation runs.
template <class SD, class R, bool B6, bool BT, int NR>
__device__ __forceinline__ void dv_cells(const FDev &F, Consts<double> K, Consts<R> KR) {
    constexpr bool NL = NR > 0;
    static_assert(!NL || (!BT && sizeof(SD) == 8 && NR <= 4), "node tables: FP64 single case");
    int i;
    bool active = true;
    NodeLds nls{};
    int nid[NL ? NR : 1];
    if constexpr (BT) {
        const int4 bl = F.cblk[cell_block(F)];
        if ((int)threadIdx.x >= bl.y) return;
        case_consts<BT>(F, bl.z, K, KR);
        i = bl.x + threadIdx.x;
    } else if constexpr (NL) {
        const int lb = cell_block(F);
        if (lb * kBlock >= F.n) return;   // a padding block (uniform)
        nls = node_lds(NR * kBlock);
        node_lds_ids<NR>(F, lb, nid);
        // lanes past the last cell compute cell n-1 and store nothing, so that
        // every lane of a wave takes part in its copies
        i = lb * kBlock + threadIdx.x;
        active = i < F.n;
        if (!active) i = F.n - 1;
    } else {
        i = cell_block(F) * kBlock + threadIdx.x;
        if (i >= F.n) return;
    }
    const int nt = F.nt, n = F.n;
    const double *Q = F.qn;
    // DT enters D only through VOL/DT: with LOCAL = 1 the record holds it, with
    // LOCAL = 0 the record holds VOL and DT is the global minimum.  NL runs with
    // LOCAL = 1 only: a DTMIN load under a branch would make the join wait for
    // every load in flight, the copies included.
    const int local = NL ? 1 : K.local;
    const double dti = local ? 0.0 : __longlong_as_double((long long)*F.dtslot);
    if (!local) F.dt[i] = dti;
    int4 nbt = F.nct[i], nd4;
    if constexpr (NL) {
        // slots in the block's LDS: the wave's region, then the slot in its list
        const int2 sl = F.nloc[i];
        const int wb = (threadIdx.x / kWave) * (NR * kWave);
        nd4 = make_int4(wb + (sl.x & 0xFFFF), wb + ((unsigned)sl.x >> 16), wb + (sl.y & 0xFFFF),
                        wb + ((unsigned)sl.y >> 16));
    } else {
        nd4 = F.nod4[i];
    }
    // raw[f]: neighbour id, with the BC type in the top bits for a ghost
    const int raw[4] = {nbt.x, nbt.y, nbt.z, nbt.w};
    const int nd[4] = {nd4.x, nd4.y, nd4.z, nd4.w};
    // ---- dissipation (solver_state.f90:183-223); ghost neighbour = parent copy
    double ci[5];
    for (int q = 0; q < 5; ++q) ci[q] = Q[q * nt + i];
    if constexpr (NL) node_lds_copy<NR>(F, nls, nid);   // after the cell's own loads
    R dsum[5];
    {
        const int nbr[4] = {raw[0] & kIdMask, raw[1] & kIdMask, raw[2] & kIdMask, raw[3] & kIdMask};
        dissipation_rec<R, SD>(F, i, nbr, ci, local, dti, dsum);
        // stored now in either form (FP32-mixed overwrites it with D + VIS
        // below): the store keeps the viscous part's loads after this point,
        // which holds the register count down
        if (active)
            for (int q = 0; q < 5; ++q) stv<SD>(F.dd, (size_t)q * n + i, dsum[q]);
    }
    if constexpr (NL) node_lds_wait();
    // ---- viscous fluxes (viscous.f90:62-70, 130-217)
    R ua, va, ta, pa;
    R rhoa = (R)ci[0];
    uvtp_from_q<R>(rhoa, (R)ci[1], (R)ci[2], (R)ci[3], ua, va, ta, pa, KR);
    R ra = fdiv((R)ci[4], rhoa);
    R xa = (R)ld<SD>(F.gxcc, i), ya = (R)ld<SD>(F.gycc, i);
    Prim<R> pi;
    bool have_pi = false;
    R vu = R(0), vv = R(0), vt = R(0), vr = R(0);
    for (int f = 0; f < 4; ++f) {
        int bnode = nd[f], dnode = nd[(f + 1) & 3], c = raw[f] & kIdMask;
        R uc, vc, tc, rhoc, c5c;
        if (raw[f] < n) {
            R pc;
            rhoc = (R)Q[c];
            uvtp_from_q<R>(rhoc, (R)Q[nt + c], (R)Q[2 * nt + c], (R)Q[3 * nt + c], uc, vc, tc, pc, KR);
            c5c = (R)Q[4 * nt + c];
        } else {
            if (!have_pi) {
                prim_from_q<R>(pi, rhoa, (R)ci[1], (R)ci[2], (R)ci[3], (R)ci[4], KR);
                have_pi = true;
            }
            // type 3: the neighbour across the opposite face, a cell
            const int ty = (unsigned)raw[f] >> kTypeShift, o = raw[(f + 2) & 3];
            R uo = R(0), vo = R(0), to = R(0), po = R(0);
            if (ty == 3)
                uvtp_from_q<R>((R)Q[o], (R)Q[nt + o], (R)Q[2 * nt + o], (R)Q[3 * nt + o], uo, vo,
                               to, po, KR);
            Prim<R> s;
            if (B6 && ty == 6) {
                bc6_load<R>(F, c - n, uc, vc, tc, rhoc, c5c);
            } else {
                ghost_state<R>(s, ty, pi, uo, vo, to, po, (R)ld<SD>(F.gycc, c), KR);
                uc = s.u; vc = s.v; tc = s.t; rhoc = s.c[0]; c5c = s.c[4];
            }
        }
        R rc = fdiv(c5c, rhoc);
        R sx, sy;
        ViscFace<R> o;
        if constexpr (NL) {
            // the node records from LDS; the face vector as face_sc forms it
            const double2 bxy = nls.xy[bnode], dxy = nls.xy[dnode];
            const double2 buv = nls.uv[bnode], duv = nls.uv[dnode];
            const double2 btr = nls.tr[bnode], dtr = nls.tr[dnode];
            const double brn = __hiloint2double((int)nls.rhi[bnode], (int)nls.rlo[bnode]);
            const double drn = __hiloint2double((int)nls.rhi[dnode], (int)nls.rlo[dnode]);
            if (F.sc_nodes == 2) {
                sx = (dxy.y - bxy.y) * (bxy.y + dxy.y) * 0.5;
                sy = -((dxy.x - bxy.x) * (bxy.y + dxy.y) * 0.5);
            } else if (F.sc_nodes == 1) {
                sx = dxy.y - bxy.y;
                sy = -(dxy.x - bxy.x);
            } else {
                face_sc<R, SD, false>(F, f, i, 0, 0, sx, sy);
            }
            o = viscous_face<R>(sx, sy, xa, ya, bxy.x, bxy.y, (R)ld<SD>(F.gxcc, c),
                                (R)ld<SD>(F.gycc, c), dxy.x, dxy.y, ua, va, ta, rhoa, ra, buv.x,
                                buv.y, btr.x, btr.y, brn, uc, vc, tc, rhoc, rc, duv.x, duv.y,
                                dtr.x, dtr.y, drn, KR);
        } else {
            face_sc<R, SD, true>(F, f, i, bnode, dnode, sx, sy);
            o = viscous_face_rec<R, SD>(F, sx, sy, xa, ya, ua, va, ta, rhoa, ra, bnode, dnode,
                                        (R)ld<SD>(F.gxcc, c), (R)ld<SD>(F.gycc, c), uc, vc, tc,
                                        rhoc, rc, KR);
        }
        vu = vu + o.u0 + o.u1;
        vv = vv + o.v0 + o.v1;
        vt = vt + o.t0 + o.t1;
        vr = vr + o.r0 + o.r1;
    }
    if (!active) return;
    if constexpr (kCombinedDV<SD>) {
        const R e[5] = {dsum[0], dsum[1] + vu, dsum[2] + vv, dsum[3] + vt, dsum[4] + vr};
        for (int q = 0; q < 5; ++q) stv<SD>(F.dd, (size_t)q * n + i, e[q]);
    } else {
        stv<SD>(F.vis, i, vu);
        stv<SD>(F.vis, (size_t)n + i, vv);
        stv<SD>(F.vis, (size_t)2 * n + i, vt);
        stv<SD>(F.vis, (size_t)3 * n + i, vr);
    }
}


template <class SD, class R, bool B6, bool BT>
__global__ void __launch_bounds__(kBlock) k_dv(FDev F, Consts<double> K, Consts<R> KR) {
    dv_cells<SD, R, B6, BT, 0>(F, K, KR);
}

// k_dv with the wave node tables (NR copy rounds): bounded to three waves per
// SIMD, the register count of k_dv itself (the compiler then spills 20 bytes per
// lane; unbounded it takes 171 registers and two waves per SIMD).
template <bool B6, int NR>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_dv_nodes(FDev F, Consts<double> K, Consts<double> KR) {
    dv_cells<double, double, B6, false, NR>(F, K, KR);
}

// ---------------------------------------------------------------------------------
// k_vface: the viscous terms of every face once (viscous.f90:130-217, evaluated
// from the owning cell's side: A = owner, C = the cell or ghost across, B and D
// the face's nodes), for the FP32-mixed stage 1 to gather.  The reference
// evaluates each interior face twice, once from each side; the two agree up to
// rounding (the gradients and the surface vector change sign together, the
// four-point averages only by summation order), so the other cell applies the
// owner's terms negated.  One lane per face, faces in owner-cell order.
template <class SD, class R, bool B6>
__global__ void __launch_bounds__(kBlock) k_vface(FDev F, Consts<R> KR) {
    const int e = (F.remap ? xcd_chunk(blockIdx.x, F.nfb8) : blockIdx.x) * kBlock + threadIdx.x;
    if (e >= F.nf) return;
    const int4 fr = F.vfl[e];
    const int i = fr.x & kIdMask, f = (unsigned)fr.x >> kTypeShift;
    const int raw = fr.y, bn = fr.z, dn = fr.w, c = raw & kIdMask;
    const int n = F.n, nt = F.nt;
    const double *Q = F.qn;
    const R rhoa = (R)Q[i], c1a = (R)Q[nt + i], c2a = (R)Q[2 * nt + i], c3a = (R)Q[3 * nt + i],
            c5a = (R)Q[4 * nt + i];
    R ua, va, ta, pa;
    uvtp_from_q<R>(rhoa, c1a, c2a, c3a, ua, va, ta, pa, KR);
    R uc, vc, tc, rhoc, c5c;
    if (raw < n) {
        R pc;
        rhoc = (R)Q[c];
        uvtp_from_q<R>(rhoc, (R)Q[nt + c], (R)Q[2 * nt + c], (R)Q[3 * nt + c], uc, vc, tc, pc, KR);
        c5c = (R)Q[4 * nt + c];
    } else {
        // the ghost across this boundary face, from the owner (boundary.f90); a
        // type-3 ghost extrapolates from the neighbour across the opposite face
        Prim<R> pi, sg;
        prim_from_q<R>(pi, rhoa, c1a, c2a, c3a, c5a, KR);
        const int ty = (unsigned)raw >> kTypeShift;
        R uo = R(0), vo = R(0), to = R(0), po = R(0);
        if (ty == 3) {
            const int4 nb = F.nct[i];
            const int o = sel4((f + 2) & 3, nb.x, nb.y, nb.z, nb.w) & kIdMask;
            uvtp_from_q<R>((R)Q[o], (R)Q[nt + o], (R)Q[2 * nt + o], (R)Q[3 * nt + o], uo, vo, to,
                           po, KR);
        }
        ghost_state<R>(sg, ty, pi, uo, vo, to, po, (R)ld<SD>(F.gycc, c), KR, g6_for<B6, R>(F, c - n));
        uc = sg.u; vc = sg.v; tc = sg.t; rhoc = sg.c[0]; c5c = sg.c[4];
    }
    const ViscFace<R> o = viscous_face_rec<R, SD>(
        F, (R)ld<SD>(F.gsc, (size_t)(2 * f) * n + i), (R)ld<SD>(F.gsc, (size_t)(2 * f + 1) * n + i),
        (R)ld<SD>(F.gxcc, i), (R)ld<SD>(F.gycc, i), ua, va, ta, rhoa, fdiv(c5a, rhoa), bn, dn,
        (R)ld<SD>(F.gxcc, c), (R)ld<SD>(F.gycc, c), uc, vc, tc, rhoc, fdiv(c5c, rhoc), KR);
    F.vflux[e] = make_float4((float)(o.u0 + o.u1), (float)(o.v0 + o.v1), (float)(o.t0 + o.t1),
                             (float)(o.r0 + o.r1));
}

// ---------------------------------------------------------------------------------
// k_rk: one RK stage for one cell per lane.  Stage 1 also forms the dissipation D
// (solver_state.f90:120-223) and viscous fluxes VIS (viscous.f90:62-217) from
// Q^n and stores them for stages 2-4 (they stay frozen, as in the reference).
// The faces are visited once, in order: the neighbour's primitives (or its ghost,
// built from this cell), the face's flux terms, its viscous terms (stage 1) and
// its share of the source-term gradient differences.  SI / SO: stage input /
// output storage types; SD: storage of D, VIS, sensors and nodal averages.
// DV = 0: stage 1 reads D and VIS written by k_dv (the split variant); DV = 2
// (FP32-mixed): D inline, VIS gathered from the face terms k_vface wrote.
template <class R, class SI, class SO, class SD, int STAGE, int DV, bool B6, bool BT, bool PL>
__global__ void __launch_bounds__(kBlock) k_rk(FDev F, Consts<double> K, Consts<R> KR,
                                               double alpha) {
    constexpr bool kInlineDV = STAGE == 1 && DV != 0;
    constexpr bool kFaceVis = STAGE == 1 && DV == 2;
    const int lb = cell_block(F);
    int i, bcount = 0;
    if constexpr (BT) {
        const int4 bl = F.cblk[lb];
        case_consts<BT>(F, bl.z, K, KR);
        i = bl.x + threadIdx.x;
        bcount = bl.y;
    } else {
        i = lb * kBlock + threadIdx.x;
    }
    const int nt = F.nt, n = F.n;
    const void *qin = STAGE == 1 ? (const void *)F.qn : (const void *)F.st[STAGE & 1];
    const size_t sin = STAGE == 1 ? (size_t)nt : (size_t)n;
    double cnew[5] = {0, 0, 0, 0, 0}, cold[5] = {0, 0, 0, 0, 0};
    const bool active = BT ? (int)threadIdx.x < bcount : i < n;
    if (active) {
        // neighbour ids (ghost ids carry their BC type in the top bits)
        const int4 nbt = F.nct[i];
        const int4 nb4 = make_int4(nbt.x & kIdMask, nbt.y & kIdMask, nbt.z & kIdMask,
                                   nbt.w & kIdMask);
        const double *Q = F.qn;
        double dti;
        if (STAGE == 1) {
            dti = K.local ? F.dt[i] : __longlong_as_double((long long)*F.dtslot);
            if (!K.local) F.dt[i] = dti;
        } else {
            dti = F.dt[i];
        }
        double dd[5], vs[4] = {0.0, 0.0, 0.0, 0.0};
        R dsum[5];
        constexpr bool kPreload = PL || sizeof(R) == 4;
        // stage 1 with D inline: the dissipation's neighbour states (Q^n, the
        // stage input) double as the preloaded flux states
        constexpr bool kKeepQ = kInlineDV && kPreload;
        R qf[4][5];
        double ci[5];   // stage 1 with D inline: Q^n of the cell, which is also the
                        // stage input and COLD (loaded once for all three)
        if (kInlineDV) {
            // dissipation faces; a ghost neighbour is a copy of this cell
#pragma unroll
            for (int q = 0; q < 5; ++q) ci[q] = Q[q * nt + i];
            const int nbr[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
            if constexpr (kKeepQ)
                dissipation_rec<R, SD, true>(F, i, nbr, ci, K.local, dti, dsum, qf);
            else
                dissipation_rec<R, SD>(F, i, nbr, ci, K.local, dti, dsum);
            // FP64 stores D now.  FP32-mixed keeps it in registers and stores
            // D + VIS after the faces: its store here never runs (n >= 1), but
            // as a store it keeps the face loop's loads after this point, which
            // holds the register count down, and it writes nothing twice
#pragma unroll
            for (int q = 0; q < 5; ++q) dd[q] = (double)(SD)dsum[q];
            if (!kCombinedDV<SD> || n < 0) {
#pragma unroll
                for (int q = 0; q < 5; ++q) stv<SD>(F.dd, (size_t)q * n + i, dd[q]);
            }
        } else {
            // D and VIS frozen at stage 1 (FP32-mixed: combined in D)
#pragma unroll
            for (int q = 0; q < 5; ++q) dd[q] = ld<SD>(F.dd, (size_t)q * n + i);
            if constexpr (!kCombinedDV<SD>) {
#pragma unroll
                for (int q = 0; q < 4; ++q) vs[q] = ld<SD>(F.vis, (size_t)q * n + i);
            }
        }

        // FP32: every face's neighbour state is loaded up front (a ghost face
        // loads this cell again), so no load waits on a branch.  FP64 loads
        // them face by face unless PL: holding all twenty costs the second
        // wavefront per SIMD (256 VGPRs), which only grids too small to give a
        // SIMD two wavefronts can spare.  (Stage 1 with D inline has them from
        // the dissipation already.)
        if constexpr (!kKeepQ) {
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const int jn = sel4(f, nb4.x, nb4.y, nb4.z, nb4.w);
                const int j = jn < n ? jn : i;
#pragma unroll
                for (int q = 0; q < 5; ++q)
                    qf[f][q] = kPreload ? (R)ld<SI>(qin, q * sin + j) : R(0);
            }
        }
        Prim<R> pi;
        if constexpr (kKeepQ)
            prim_from_q<R>(pi, (R)ci[0], (R)ci[1], (R)ci[2], (R)ci[3], (R)ci[4], KR);
        else
            prim_from_q<R>(pi, (R)ld<SI>(qin, i), (R)ld<SI>(qin, sin + i),
                           (R)ld<SI>(qin, 2 * sin + i), (R)ld<SI>(qin, 3 * sin + i),
                           (R)ld<SI>(qin, 4 * sin + i), KR);
        R acc[5];
        R vu = R(0), vv = R(0), vt = R(0), vr = R(0);
        SrcNbr<R> sa{}, sb{};
        SrcDiff<R> sd{};
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const int jn = sel4(f, nb4.x, nb4.y, nb4.z, nb4.w);
            const int ty = (unsigned)sel4(f, nbt.x, nbt.y, nbt.z, nbt.w) >> kTypeShift;
            Prim<R> pj;
            if (ty == 0) {
                if (kPreload)
                    prim_from_q<R>(pj, qf[f][0], qf[f][1], qf[f][2], qf[f][3], qf[f][4], KR);
                else
                    prim_from_q<R>(pj, (R)ld<SI>(qin, jn), (R)ld<SI>(qin, sin + jn),
                                   (R)ld<SI>(qin, 2 * sin + jn), (R)ld<SI>(qin, 3 * sin + jn),
                                   (R)ld<SI>(qin, 4 * sin + jn), KR);
            } else {
                // the ghost across this face: only this cell reads it (boundary.f90);
                // a type-3 ghost extrapolates from the neighbour across the
                // opposite face, which is a cell (checked when the mesh is set up)
                const int fo = (f + 2) & 3;
                R uo = R(0), vo = R(0), to = R(0), po = R(0);
                if (ty == 3) {
                    if (kPreload) {
                        uvtp_from_q<R>(qf[fo][0], qf[fo][1], qf[fo][2], qf[fo][3], uo, vo, to, po,
                                       KR);
                    } else {
                        const int o = sel4(fo, nb4.x, nb4.y, nb4.z, nb4.w);
                        uvtp_from_q<R>((R)ld<SI>(qin, o), (R)ld<SI>(qin, sin + o),
                                       (R)ld<SI>(qin, 2 * sin + o), (R)ld<SI>(qin, 3 * sin + o),
                                       uo, vo, to, po, KR);
                    }
                }
                ghost_state<R>(pj, ty, pi, uo, vo, to, po, (R)ld<SD>(F.gycc, jn), KR,
                               g6_for<B6, R>(F, jn - n));
            }
            // central flux terms of this face (residual.f90:34-54)
            const R sx = (R)ld<SD>(F.gsc, (2 * f) * n + i);
            const R sy = (R)ld<SD>(F.gsc, (2 * f + 1) * n + i);
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                const R tf = (pi.f[q] + pj.f[q]) * sx;
                const R tg = (pi.g[q] + pj.g[q]) * sy;
                acc[q] = f == 0 ? tf + tg : acc[q] + tf + tg;
            }
            const R xj = (R)ld<SD>(F.gxcc, jn), yj = (R)ld<SD>(F.gycc, jn);
            const SrcNbr<R> nbv = src_nbr<R>(xj, yj, pj.u, pj.v, pj.c[0], pj.c[4]);
            if (f == 0) sa = nbv;
            if (f == 1) sb = nbv;
            if (f == 2) src_pair(sa, nbv, sd.xac, sd.yac, sd.qac, sd.rac, sd.uac, sd.vac);
            if (f == 3) src_pair(sb, nbv, sd.xbd, sd.ybd, sd.qbd, sd.rbd, sd.ubd, sd.vbd);
            if (kInlineDV && !kFaceVis) {
                const int4 nd4 = F.nod4[i];
                const int bn = sel4(f, nd4.x, nd4.y, nd4.z, nd4.w);
                const int dn = sel4(f, nd4.y, nd4.z, nd4.w, nd4.x);
                const ViscFace<R> o = viscous_face_rec<R, SD>(
                    F, sx, sy, (R)ld<SD>(F.gxcc, i), (R)ld<SD>(F.gycc, i), pi.u, pi.v, pi.t,
                    pi.c[0], fdiv(pi.c[4], pi.c[0]), bn, dn, xj, yj, pj.u, pj.v, pj.t, pj.c[0],
                    fdiv(pj.c[4], pj.c[0]), KR);
                vu = vu + o.u0 + o.u1;
                vv = vv + o.v0 + o.v1;
                vt = vt + o.t0 + o.t1;
                vr = vr + o.r0 + o.r1;
            }
        }
        if (kFaceVis) {
            float4 vfl4[4];
            int vfsg = 0;
            const int4 cf = F.cvf[i];
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const int ef = sel4(f, cf.x, cf.y, cf.z, cf.w);
                vfl4[f] = F.vflux[ef & 0x7fffffff];
                vfsg |= ef < 0 ? 1 << f : 0;
            }
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const R sg = (vfsg >> f) & 1 ? R(-1) : R(1);
                vu = vu + sg * (R)vfl4[f].x;
                vv = vv + sg * (R)vfl4[f].y;
                vt = vt + sg * (R)vfl4[f].z;
                vr = vr + sg * (R)vfl4[f].w;
            }
        }
        if (kInlineDV) {
            if constexpr (kCombinedDV<SD>) {
                const R e[5] = {dsum[0], dsum[1] + vu, dsum[2] + vv, dsum[3] + vt, dsum[4] + vr};
#pragma unroll
                for (int q = 0; q < 5; ++q) {
                    dd[q] = (double)(SD)e[q];
                    stv<SD>(F.dd, (size_t)q * n + i, dd[q]);
                }
            } else {
                vs[0] = (double)(SD)vu; vs[1] = (double)(SD)vv;
                vs[2] = (double)(SD)vt; vs[3] = (double)(SD)vr;
#pragma unroll
                for (int q = 0; q < 4; ++q) stv<SD>(F.vis, (size_t)q * n + i, vs[q]);
            }
        }
        R s3;
        const R s5 = source_terms_d<R>(sd, pi.c[0], pi.c[4], pi.t, pi.p, pi.v,
                                       (R)ld<SD>(F.gycc, i), s3, KR);
        // residual.f90:34-60 and the update residual.f90:76-81 (FP64)
        const double vol = F.vol[i];
        double r[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) r[q] = (double)(R(0.5) * acc[q]);
        r[4] = r[4] - vol * (double)s5;
        if (K.iaxis == 1) r[2] = r[2] - vol * (double)s3;
        const double a = alpha * fdiv(dti, vol);
#pragma unroll
        for (int q = 0; q < 5; ++q) cold[q] = kKeepQ ? ci[q] : Q[q * nt + i];
        cnew[0] = cold[0] - a * (r[0] - dd[0]);
        cnew[1] = cold[1] - a * (r[1] - vs[0] - dd[1]);
        cnew[2] = cold[2] - a * (r[2] - vs[1] - dd[2]);
        cnew[3] = cold[3] - a * (r[3] - vs[2] - dd[3]);
        cnew[4] = cold[4] - a * kCflTurb * (r[4] - vs[3] - dd[4]);
        cnew[4] = cnew[0] * dmax(K.rtol, fdiv(cnew[4], cnew[0]));
        if (STAGE < 4) {
            void *qo = F.st[(STAGE - 1) & 1];
#pragma unroll
            for (int q = 0; q < 5; ++q) stv<SO>(qo, (size_t)q * n + i, cnew[q]);
        } else {
            double *Qw = F.qn;
#pragma unroll
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
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const double d = cnew[q] - cold[q];
            sq[q] = active ? d * d : 0.0;
            mx[q] = active ? fabs(d) : 0.0;
            bad |= active && !isfinite(cnew[q]);
        }
        // slot = logical block: the same cells, summed in the same order, as the
        // routine-level monitor's block partials
        if (BT ? bcount > 0 : lb * kBlock < n)
            block_partials_write(sq, mx, bad, F.part + (size_t)lb * kPartStride);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            // batch: this iteration's pre-pass has folded the previous partials
            if (BT && F.ctr[3]) F.ctr[0] += 1;
            F.ctr[3] = 1;                                    // partials pending
        }
    }
}

}  // namespace
}  // namespace tdt

#include "quad_kernels.hpp"

namespace tdt {
namespace {



// Launches one kernel; when profiling, with start/stop events that the runtime
// stamps at the kernel's own begin and end (same durations a kernel trace shows).
template <class K, class... Args>
void launch_shm(tdt_ctx *c, FusedBufs &B, int slot, K kern, unsigned grid, size_t shm,
                Args... args) {
    if (c->profiling && B.have_events)
        hipExtLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), shm, c->stream, B.ev[2 * slot],
                              B.ev[2 * slot + 1], 0, args...);
    else
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), shm, c->stream, args...);
}

template <class K, class... Args>
void launch(tdt_ctx *c, FusedBufs &B, int slot, K kern, unsigned grid, Args... args) {
    launch_shm(c, B, slot, kern, grid, 0, args...);
}

// k_dv with the wave node tables can run: FP64 single case (the tables exist)
// with LOCAL = 1.
inline bool node_lds_usable(const tdt_ctx *c, const FusedBufs &B) {
    return B.F.nl_cap > 0 && B.F.ncases == 0 && c->K.local;
}

// k_dv_nodes for the table's rounds per wave (nl_cap / 64, 1..4).
template <bool B6>
void launch_dv_nodes(tdt_ctx *c, FusedBufs &B) {
    const FDev &F = B.F;
    const size_t shm = (size_t)F.nl_cap * (kBlock / kWave) * kNodeLdsBytes;
    switch (F.nl_cap / kWave) {
    case 1: launch_shm(c, B, 1, k_dv_nodes<B6, 1>, F.ncb8, shm, F, c->K, c->K); break;
    case 2: launch_shm(c, B, 1, k_dv_nodes<B6, 2>, F.ncb8, shm, F, c->K, c->K); break;
    case 3: launch_shm(c, B, 1, k_dv_nodes<B6, 3>, F.ncb8, shm, F, c->K, c->K); break;
    default: launch_shm(c, B, 1, k_dv_nodes<B6, 4>, F.ncb8, shm, F, c->K, c->K);
    }
}

// One-lane stage kernel, with the neighbour states preloaded or not (FP64; FP32
// always preloads).
template <class R, class SI, class SO, class SD, int STAGE, int DV, bool B6, bool BT>
void launch_rk(tdt_ctx *c, FusedBufs &B, int slot, bool preload, const Consts<R> &KR,
               double alpha) {
    if constexpr (sizeof(R) == 8) {
        if (preload) {
            launch(c, B, slot, k_rk<R, SI, SO, SD, STAGE, DV, B6, BT, true>, B.F.ncb8, B.F,
                   c->K, KR, alpha);
            return;
        }
    }
    launch(c, B, slot, k_rk<R, SI, SO, SD, STAGE, DV, B6, BT, false>, B.F.ncb8, B.F, c->K, KR,
           alpha);
}

template <class R, class SS, class SD, bool B6, bool BT>
hipError_t enqueue_iteration(tdt_ctx *c, FusedBufs &B, const Consts<R> &KR) {
    const FDev &F = B.F;
    const int ks = BT ? B.kset & (kSt1Split | kStPreload) : B.kset;   // a batch: no quad form
    const bool pl = ks & kStPreload;
    const unsigned nbc = F.ncb8, nq = B.nbq_cells;
    const int nparts = partial_blocks(B, ks);
    auto no_dv_slot = [&]() {   // slot 1 unused: dissipation and viscous live in stage 1
        if (c->profiling && B.have_events) {
            (void)hipEventRecord(B.ev[2], c->stream);
            (void)hipEventRecord(B.ev[3], c->stream);
        }
    };
    // type-6 ghosts of each stage's input, ahead of the kernels that read them
    const unsigned ngb = (unsigned)((F.g + kBlock - 1) / kBlock);
    auto bc6 = [&](auto kern) {
        if (B6) hipLaunchKernelGGL(kern, dim3(ngb), dim3(kBlock), 0, c->stream, F, KR);
    };
    bc6(k_bc6<R, double, SD, 1>);
    if (ks & kPreQuad) {
        launch(c, B, 0, kq_pre<SD, R, B6>, nq + B.nbq_nodes + 1, F, c->K, KR, B.nbq_cells,
               nparts);
    } else {
        const int nnb8 = (B.nb_nodes + 7) & ~7;
        const unsigned nfold8 = ((BT ? F.ncases : 1) + 7) & ~7;   // the folding blocks, padded
        launch(c, B, 0, k_pre<SD, R, B6, BT>, nfold8 + F.ncb8 + nnb8, F, c->K, KR, nparts,
               F.ncb8, nnb8);
    }
    switch (ks & kSt1Mask) {
    case kSt1Split:
        if constexpr (sizeof(SD) == 8 && !BT) {
            if ((ks & kDvNodeLds) && node_lds_usable(c, B)) {
                launch_dv_nodes<B6>(c, B);
                launch_rk<R, double, SS, SD, 1, false, B6, BT>(c, B, 2, pl, KR, 0.25);
                break;
            }
        }
        launch(c, B, 1, k_dv<SD, R, B6, BT>, nbc, F, c->K, KR);
        launch_rk<R, double, SS, SD, 1, false, B6, BT>(c, B, 2, pl, KR, 0.25);
        break;
    case kSt1Quad:
        no_dv_slot();
        launch(c, B, 2, kq_stage<R, double, SS, SD, 1, B6>, nq, F, c->K, KR, 0.25);
        break;
    case kSt1Face:   // FP32-mixed single cases only (fused_setup refuses it otherwise)
        if constexpr (sizeof(R) == 4 && !BT) {
            launch(c, B, 1, k_vface<SD, R, B6>, F.nfb8, F, KR);
            launch(c, B, 2, k_rk<R, double, SS, SD, 1, 2, B6, BT, false>, nbc, F, c->K, KR, 0.25);
            break;
        }
        [[fallthrough]];
    default:
        no_dv_slot();
        launch(c, B, 2, k_rk<R, double, SS, SD, 1, true, B6, BT, false>, nbc, F, c->K, KR, 0.25);
    }
    if (ks & kStQuad) {
        bc6(k_bc6<R, SS, SD, 2>);
        launch(c, B, 3, kq_stage<R, SS, SS, SD, 2, B6>, nq, F, c->K, KR, 1.0 / 3.0);
        bc6(k_bc6<R, SS, SD, 3>);
        launch(c, B, 4, kq_stage<R, SS, SS, SD, 3, B6>, nq, F, c->K, KR, 0.5);
        bc6(k_bc6<R, SS, SD, 4>);
        launch(c, B, 5, kq_stage<R, SS, double, SD, 4, B6>, nq, F, c->K, KR, 1.0);
    } else {
        bc6(k_bc6<R, SS, SD, 2>);
        launch_rk<R, SS, SS, SD, 2, true, B6, BT>(c, B, 3, pl, KR, 1.0 / 3.0);
        bc6(k_bc6<R, SS, SD, 3>);
        launch_rk<R, SS, SS, SD, 3, true, B6, BT>(c, B, 4, pl, KR, 0.5);
        bc6(k_bc6<R, SS, SD, 4>);
        launch_rk<R, SS, double, SD, 4, true, B6, BT>(c, B, 5, pl, KR, 1.0);
    }
    return hipGetLastError();
}

}  // namespace
}  // namespace tdt
