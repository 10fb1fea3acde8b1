// Shared definitions of the fused hot path: device view, host-side buffers
// and the per-precision enqueue functions (fused.hip: FP64, strict IEEE;
// fused_mixed.hip: FP32-mixed, compiled with fast reciprocal / approximate
// library functions).
#pragma once

#include "ctx.hpp"
#include "reduce.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdlib>
#include <string>
#include <vector>

namespace tdt {

// Neighbour tables of the fused kernels carry a ghost's boundary type in the
// top bits of its id, so a face's type needs no second dependent load.
constexpr int kTypeShift = 28;
constexpr int kIdMask = (1 << kTypeShift) - 1;
// Node gathers of up to this many contributors use fixed-width records.
constexpr int kNodeWidth = 4;
// Values per cell record (drec) and their slots.
constexpr int kRecLen = 8;
constexpr int kRecAnu = 0, kRecAnurt = 1, kRecVbydt = 2, kRecD2q = 3;

struct FDev {
    int n, g, m, nt;
    int remap;                      // XCD-aware block order
    int ncb8;                       // cell blocks rounded up to a multiple of 8
    const int4 *nc4, *nod4;
    const int4 *nct;                // [n] nc4 with ghost ids | type << kTypeShift
    const int2 *ncon;               // [m][kNodeWidth] contributors {a, source cell}: a =
                                    //   cell id, ghost id | type << 28, or for a
                                    //   type-3 ghost its opposite cell | 3 << 28
    const int *nrel;                // [m] count | U, V, nu~ sum starts << 4, 8, 12 |
                                    //   nu~ floor flag << 16
    const double *xn, *yn;
    const int *gparent, *gtype, *gopp, *gside;
    const int *g6ptr, *g6cell;      // BC type 6 port columns
    const double *g6dely;
    int has_bc6;                    // any type-6 ghost
    void *g6s;                      // [6][g] type-6 ghosts of the current stage (k_bc6)
    const int *nptr, *nlist;
    const int4 *novr;
    const double *xcc, *ycc, *vol, *dl, *sc;
    // geometry as the flux arithmetic reads it, in the storage type SD: the FP64
    // arrays themselves, or FP32 copies rounded from them (FP32-mixed)
    const void *gxcc, *gycc, *gvol, *gsc, *gxn, *gyn;
    int sc_nodes;                   // FP64: the SC arrays are the node coordinates' face
                                    // vectors (1 planar, 2 axisymmetric), so kernels that
                                    // gather the nodes may form them; 0: read SC
    double *qn;                     // [5][nt] Q^n (the C array), FP64
    void *st[2];                    // [5][n] stage states Q(1), Q(2)/Q(3)
    double *dt;                     // [n]
    // drec[n]: per-iteration cell records of eight values of the storage type
    // SD, ANU, ANURT, VOL/DT (VOL when LOCAL = 0), D2Q1..D2Q5 of Q^n, in planes
    // of 16-byte vectors (a gather of one costs two or four 16-byte loads)
    void *drec;
    // nodal averages UN, VN, TN, RHON as 16-byte vectors (one plane in FP32,
    // two in FP64), then RN: 5 * m values of SD; nxy[m]: the node coordinates
    // as SD pairs (float2 / double2), refreshed when the geometry changes
    void *nrec, *nxy;
    void *dd;                       // [5][n] dissipation D
    void *vis;                      // [4][n]
    unsigned long long *dtslot;     // DTMIN bits for LOCAL = 0
    double *part;                   // per-block monitor partials
    double *hist;                   // [niter][10]
    int *ctr;                       // 0: iteration, 2: first non-finite iteration,
                                    // 3: stage-4 monitor partials pending
    // batch of independent cases (ncases > 0): each logical cell / node block
    // holds cells / nodes of one case, {first, count, case, 0}; ccase[c] =
    // {first logical cell block, cell blocks, cells, 0}; kc64 / kcr: the cases'
    // flow constants (Consts<double>, Consts<R>)
    int ncases;
    const int4 *cblk, *nblk, *ccase;
    const void *kc64, *kcr;
    // viscous faces (FP32-mixed kernel sets with kSt1Face): every face once,
    // vfl[e] = {owner cell | face << kTypeShift, neighbour as in nct, node b,
    // node d} (an interior face belongs to the lower cell id, a boundary face to
    // its cell); vflux[e] = the face's VISU, VISV, VIST, VISR terms from the
    // owner's side; cvf[i] = cell i's four face indices, bit 31 set where i is
    // not the owner (it applies the terms negated)
    int nf, nfb8;
    const int4 *vfl;
    float4 *vflux;
    const int4 *cvf;
    // node tables of the wavefronts (FP64 single case, kDvNodeLds): wave wg (64
    // cells of logical block wg / 4) lists the distinct nodes of its cells in
    // bnode[wg * nl_cap ..], ascending and padded with the last to nl_cap (a
    // multiple of 64, at most 256) entries; nloc[i] = cell i's four nodes' slots
    // in its wave's list, 16 bits each: {slot0 | slot1 << 16, slot2 | slot3 << 16};
    // nl_cap = 0: no tables
    const int *bnode;
    const int2 *nloc;
    int nl_cap;
};

struct FusedBufs {
    FDev F{};
    int nb_cells = 0, nb_nodes = 0;
    size_t hist_cap = 0;
    double *hist_host = nullptr;    // pinned copy target of the history, hist_cap records
    std::vector<void *> own;
    hipGraph_t graph1 = nullptr, graphK = nullptr;
    hipGraphExec_t exec1 = nullptr, execK = nullptr;
    int chunk = 8;
    hipEvent_t ev[12];
    bool have_events = false;
    int kset = 0;                    // kernel set of the iteration (kPre* | kSt1* | kSt*)
    bool tuned = false;              // kernel set picked by timing (autotune)
    double *qbak = nullptr;          // Q^n saved while candidates are timed
    int nbq_cells = 0, nbq_nodes = 0;
    int geom_version = -1;           // tdt_ctx::geom_version the FP32 copies match
};

// Kernel set of one iteration, chosen per slot; every choice computes the same
// values (FP64: bitwise).  Bit 0: pre-pass; bits 1-2: stage 1; bit 3: stages 2-4;
// bit 4: how the one-lane stage kernels load neighbour states.
constexpr int kPreQuad = 1;        // kq_pre instead of k_pre
constexpr int kSt1Split = 1 << 1;  // k_dv + k_rk<1, DV=false> instead of k_rk<1> (merged)
constexpr int kSt1Quad = 2 << 1;   // kq_stage<1>
constexpr int kSt1Face = 3 << 1;  // FP32-mixed, single case: k_vface (each viscous face
                                   // once) + k_rk<1> with D inline and VIS gathered
constexpr int kSt1Mask = 3 << 1;
constexpr int kStQuad = 1 << 3;    // kq_stage<2..4> instead of k_rk<2..4>
// FP64 one-lane k_rk (split stage 1, stages 2-4): all four neighbours' states
// requested up front instead of face by face.  More registers (one wavefront
// per SIMD), one memory round trip instead of four: pays when the grid is too
// small to give a SIMD more than one wavefront anyway.  FP32 always preloads.
constexpr int kStPreload = 1 << 4;
// FP64 k_dv (split stage 1, single case, LOCAL = 1): each wavefront copies the
// node records of its cells into LDS with direct global-to-LDS loads, which hold
// no registers in flight, and the viscous faces read them there instead of
// gathering them face by face.
constexpr int kDvNodeLds = 1 << 5;
constexpr int kSetBits = kPreQuad | kSt1Mask | kStQuad | kStPreload | kDvNodeLds;
// the uniform sets
constexpr int kSetMerged = 0;
constexpr int kSetSplit = kSt1Split;
constexpr int kSetQuad = kPreQuad | kSt1Quad | kStQuad;
constexpr int kSetFace = kSt1Face;
// stage 4 writes the monitor partials: one per block of its kind
inline int partial_blocks(const FusedBufs &B, int kset) {
    return (kset & kStQuad) ? B.nbq_cells : B.nb_cells;
}
// trial iterations per kernel set when the first tdt_iterate call picks one
constexpr int kTuneIterations = 4;


// One iteration's kernels enqueued on the context's stream (FP64 / FP32-mixed).
hipError_t enqueue_fp64(tdt_ctx *c, FusedBufs &B);
hipError_t enqueue_mixed(tdt_ctx *c, FusedBufs &B);

}  // namespace tdt
