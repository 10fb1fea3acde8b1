// Library context and device views shared by the routine-level kernels
// (routines.hip), the fused hot path (fused.hip) and the C ABI (api.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mesh_prep.hpp"
#include "phys.hpp"
#include "tdt.h"

namespace tdt {

constexpr int kBlock = 256;
// Fused kernels per iteration: pre-pass, dissipation+viscous, stages 1..4.
constexpr int kNumFusedKernels = 6;

// Device view of the mesh, geometry and routine-level (reference-layout) arrays.
// Internal ids: cells 0..n-1 (renumbered), ghosts n..n+g-1 (bc.in order),
// nodes 0..m-1.  Component-major storage: a[k*stride + id].
struct Dev {
    int n, g, m, nt;          // nt = n + g
    const int4 *nc4;          // [n] neighbours across faces 1..4
    const int4 *nod4;         // [n] nodes 1..4
    const double *xn, *yn;    // [m] metres
    const int *gparent, *gside, *gtype, *gopp;  // [g]
    const int *nptr, *nlist;  // node gather lists
    const int4 *novr;         // [m]
    const int *g6ptr, *g6cell;  // BC type 6 port columns (mesh_prep.hpp)
    const double *g6dely;
    double *xcc, *ycc, *vol;  // [nt]
    double *dl;               // [n]
    double *sc;               // [8][n] SC1X SC1Y SC2X SC2Y SC3X SC3Y SC4X SC4Y
    double *c;                // [5][nt]  C1..C5
    double *cold;             // [5][n]
    double *u, *v, *t, *p, *dt;  // [nt]
    double *f, *gf;           // [5][nt]  F1..F5, G1..G5
    double *r;                // [5][n]
    double *d2q;              // [5][nt]
    double *d24;              // [5][n]
    double *anu, *anurt;      // [nt]
    double *vis;              // [4][n]  VISU VISV VIST VISR
    double *src3, *src5;      // [n]
    double *nodeavg;          // [5][m]  UN VN TN RHON RN
    double *scal;             // [8]  0: DTMIN (as bits for atomic min) 1: RTOL ...
    double *part;             // monitor partials
    int *flags;               // [4]  0: non-finite flag
};

struct FusedBufs;  // fused.hip

// A batch of independent cases in one context (tdt_batch_create).  Cells of
// case c are internal ids [cell_off[c], cell_off[c] + ncell[c]), in the order
// tdt_create would give that case alone; its ghosts and its grid.in ids follow
// the same pattern.  Cell and node blocks never straddle two cases.
struct BatchInfo {
    int ncases = 0;
    std::vector<int> cell_off, ncell, ghost_off, nghost;
    std::vector<int> orig_cell_off;          // first combined original cell id of each case
    std::vector<tdt_params> params;
    std::vector<double> rinit;
    std::vector<Consts<double>> k64;
    std::vector<Consts<float>> kf;
    std::vector<int4> cblk, nblk, ccase;     // cblk / nblk padded to multiples of 8 blocks
    int nb_cells = 0, nb_nodes = 0;          // blocks holding cells / nodes
    Consts<double> *d_k64 = nullptr;
    Consts<float> *d_kf = nullptr;
    int4 *d_cblk = nullptr, *d_nblk = nullptr, *d_ccase = nullptr;
};

// Buffers of the residual adjoint (adjoint.hip), internal cell order: the
// state, the weight and the results, the per-cell Jacobian blocks, and the
// transposed stencil (for cell k, the (cell i, member s) whose member s is k).
struct AdjBufs {
    double *q = nullptr, *v = nullptr, *g = nullptr, *r = nullptr;
    double *J = nullptr, *Jp = nullptr, *part = nullptr, *gp = nullptr;
    int *tptr = nullptr;
    int2 *tlist = nullptr;
    // node coordinates: per-cell blocks over its 20 node slots, the per-node
    // (cell, slot) lists and the result [2][m]
    double *Jx = nullptr, *gxy = nullptr;
    int *nptr = nullptr;
    int2 *nlist = nullptr;
};

// Buffers of the iteration adjoint (adjoint_iter.hip, api_adjoint_iter.cpp):
// per-cell member lists of the dissipation / viscous / timestep operators and
// their transpose, the per-member Jacobian blocks, the stage states and
// residuals of the iteration being reversed, the adjoint work vectors, and the
// checkpoints of the unrolled forward run.
struct IterAdjBufs {
    int *mptr = nullptr, *mlist = nullptr;   // [n + 1], [members]
    int *tptr = nullptr;                     // [n + 1]
    int2 *tlist = nullptr;                   // (cell i, member entry e) by member cell
    int kmax = 0;                            // largest member count
    bool xy_ok = true;                       // node derivatives available on this mesh
    double *JE = nullptr, *JEp = nullptr, *JEx = nullptr;
    double *qn = nullptr, *qs[3] = {nullptr, nullptr, nullptr}, *r[4] = {};
    double *mu = nullptr, *lam = nullptr, *wr = nullptr, *W = nullptr;
    double *gdt = nullptr, *gvol = nullptr, *grtol = nullptr, *g = nullptr;
    double *part = nullptr, *gp = nullptr, *gxy = nullptr;   // gp: P0, T0, RTOL
    // unrolled run: checkpoints every `stride` iterations, one segment of states,
    // the final state
    // fixed-point adjoint: the residual blocks at the four stage states of the
    // fixed state, kept for the whole iteration
    double *Jst[4] = {}, *Jpst[4] = {};
    double *gJ = nullptr, *gpx = nullptr;    // dJ/dQ at the fixed state; parameter scratch
    std::vector<double *> ck;
    std::vector<double *> seg;
    double *qlast = nullptr;
    int niter = 0, stride = 0;
    bool cold = false;
    double rinit = 0.0;
};

}  // namespace tdt

struct tdt_ctx {
    std::string err;
    int precision = TDT_FP64;
    int device = 0;
    tdt_params params{};
    tdt::PreparedMesh pm;
    tdt::Consts<double> K{};
    tdt::Consts<float> Kf{};
    double rinit = 0.0;
    double params_rtred = 0.0;
    hipStream_t stream = nullptr;
    tdt::Dev D{};
    std::vector<void *> allocs;
    tdt::FusedBufs *fused = nullptr;
    int profiling = 0;
    double ktimes[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int geom_version = 0;     // bumped when tdt_set_array changes geometry
    bool sc_from_nodes = false;   // SC holds the node coordinates' face vectors (set by
                                  // the geometry pass, cleared when tdt_set_array writes SC)
    double *post = nullptr;   // [6][m] POSTPROCESS output, allocated on first use
    tdt::BatchInfo *batch = nullptr;   // tdt_batch_create contexts only
    tdt::AdjBufs *adj = nullptr;       // residual adjoint buffers, allocated on first use
    tdt::IterAdjBufs *iadj = nullptr;  // iteration adjoint buffers, allocated on first use
};

namespace tdt {

// Launch helpers implemented in routines.hip.
hipError_t launch_geometry(tdt_ctx *c, int iaxis);
hipError_t launch_timestep(tdt_ctx *c);
hipError_t launch_dissipation(tdt_ctx *c);
hipError_t launch_primitives(tdt_ctx *c);
hipError_t launch_boundary(tdt_ctx *c);
hipError_t launch_viscous(tdt_ctx *c);
hipError_t launch_source(tdt_ctx *c);
hipError_t launch_residual(tdt_ctx *c);
hipError_t launch_update(tdt_ctx *c, double alpha);
// L2 (out[0..4]) and Linf (out[5..9]) of C - COLD, flags[0] |= non-finite.
hipError_t launch_monitor(tdt_ctx *c, int save_old, double *out);
hipError_t launch_cold_start(tdt_ctx *c, double rho, double ru, double rv, double re,
                             double rr);
hipError_t launch_copy_cold(tdt_ctx *c);
// tec_plot.dat node columns U, V, P [bar], RHO, MACH, T into out[6][m].
hipError_t launch_postprocess(tdt_ctx *c, double *out);

// Residual and its adjoint (adjoint.hip); qs, v, r, g: [5][n] internal order.
hipError_t launch_res_eval(tdt_ctx *c, const double *qs, double *r);
// ... in two steps: the per-cell blocks at a state, then their transpose applied to v
hipError_t launch_res_jac(tdt_ctx *c, const double *qs, double *J, double *Jp, double *Jx);
hipError_t launch_res_apply(tdt_ctx *c, const double *J, const double *Jp, const double *v,
                            const int *tptr, const int2 *tlist, double *part, double *g,
                            double *gp);
hipError_t launch_res_vjp(tdt_ctx *c, const double *qs, const double *v, const int *tptr,
                          const int2 *tlist, double *J, double *Jp, double *part, double *g,
                          double *gp, double *Jx, const int *nptr, const int2 *nlist,
                          double *gxy);

// Iteration adjoint (adjoint_iter.hip); states [5][n] internal order.
// od: [D1..D5][n], ov: [VIS1..VIS4, DT][n]
hipError_t launch_op_eval(tdt_ctx *c, const double *qs, double *od, double *ov);
hipError_t launch_op_jac(tdt_ctx *c, const double *qs, const int *mptr, const int *mlist,
                         int kmax, double *JE, double *JEp, double *JEx);
hipError_t launch_op_vjp(tdt_ctx *c, const double *JE, const double *JEp, const double *JEx,
                         const double *W, const double *gdt, const double *gvol,
                         const double *grtol, const int *tptr, const int2 *tlist,
                         const int *nptr, const int2 *nlist, double *part, double *g,
                         double *gp, double *gxy);
hipError_t launch_stage_vjp(tdt_ctx *c, double alpha, const double *q, const double *r,
                            const double *vis, const double *d24, const double *dt,
                            const double *vol, const double *mu, double *lam, double *wr,
                            double *W, double *gdt, double *gvol, double *grtol);
hipError_t launch_add(tdt_ctx *c, int count, const double *x, double *y);
// p[0..count) = 0 by a kernel (unlike hipMemsetAsync, safe inside a captured graph);
// a[0..5n) = b[0..5n) = d[0..n) = 0 in one launch
hipError_t launch_zero(tdt_ctx *c, int count, double *p);
hipError_t launch_zero_5n_5n_n(tdt_ctx *c, double *a, double *b, double *d);
// y += x1, then += x2, then += x3 (one launch, the sums of three launch_add calls)
hipError_t launch_add3(tdt_ctx *c, int count, const double *x1, const double *x2,
                       const double *x3, double *y);
// the cells' C rows [5][nt] -> [5][n]; and [5][n] -> C and COLD
hipError_t launch_pack(tdt_ctx *c, const double *src, double *dst);
hipError_t launch_unpack(tdt_ctx *c, const double *src);

// Fused path (fused.hip).
int fused_setup(tdt_ctx *c);
void fused_free(tdt_ctx *c);
void fused_drop_graphs(tdt_ctx *c);   // constants changed: recapture on the next iterate
int fused_iterate(tdt_ctx *c, int niter, double *l2, double *linf, int *first_nonfinite);
int fused_kernel_set(tdt_ctx *c);   // kernel set in use, -1 before the first iterate

inline unsigned nblocks(int count) { return (unsigned)((count + kBlock - 1) / kBlock); }

// Sets the message tdt_last_error(NULL) returns on this thread (failures with no
// context: tdt_create, tdt_monitor_noalloc).
void api_set_error(const std::string &msg);

}  // namespace tdt
