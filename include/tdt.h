/*
 * tdt.h -- C ABI of libtdt, the MI355X (gfx950) residual + Jameson RK4 engine
 * for the 2-D / axisymmetric turbulent Navier-Stokes solver.
 *
 * The reference has no library interface: its hot path is a set of
 * argument-less Fortran subroutines sharing module state.  Each entry point
 * below replaces one of them (file:line of the reference routine given), so a
 * Fortran host can keep its CALL sites and bind these through iso_c_binding.
 *
 * Conventions
 *   - every function returns TDT_OK (0) or a negative TDT_ERR_* code;
 *     tdt_last_error(ctx) gives the message (ctx may be NULL for failures of
 *     tdt_create);
 *   - ids are 1-based as in grid.in / bc.in; arrays are host memory, FP64 or
 *     int32, contiguous; per-cell arrays with several components are
 *     column-major [component][neles] (Fortran C1..C5(NELES) / NOD(NELES,4));
 *   - the library renumbers cells and nodes internally; every array crossing
 *     this ABI is in the caller's (original) order;
 *   - the conservative state Q = [rho, rho u, rho v, rho E, rho nu~] stays in
 *     device memory between tdt_set_state / tdt_cold_start and
 *     tdt_get_state;
 *   - a context is used by one host thread at a time; contexts are
 *     independent.
 */
#ifndef TDT_H
#define TDT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDT_OK 0
#define TDT_ERR_ARG (-1)         /* bad argument or array length            */
#define TDT_ERR_MESH (-2)        /* inconsistent grid.in / bc.in data       */
#define TDT_ERR_DEVICE (-3)      /* HIP runtime error or no usable GPU      */
#define TDT_ERR_NONFINITE (-4)   /* NaN/Inf appeared in the state           */
#define TDT_ERR_UNSUPPORTED (-5) /* feature outside this library's scope    */

#define TDT_FP64 0        /* FP64 everywhere (parity path)                          */
#define TDT_FP32_MIXED 1  /* FP32 stage storage + flux/dissipation/viscous/source
                             arithmetic; FP64 geometry, Q^n, update and norms       */

#define TDT_ORDER_INPUT 0 /* keep grid.in cell order                       */
#define TDT_ORDER_RCM 1   /* reverse Cuthill-McKee                         */
#define TDT_ORDER_AUTO 2  /* TDT_ORDER_BOUNDARY_LAST on grids of 262144 cells
                           * or more, else whichever of input and RCM has
                           * the smaller bandwidth                          */
#define TDT_ORDER_BOUNDARY_LAST 3 /* the smaller-bandwidth order, then cells
                           * with a ghost neighbour moved last, grouped by
                           * the boundary type on each face                */

typedef struct tdt_ctx tdt_ctx;

/* Grid and boundary lists as read by COMPUTE_GEOMETRY (geometry.f90:31-52). */
typedef struct tdt_mesh {
    int32_t nods, neles, nghosts;
    const double *x_mm, *y_mm;   /* [nods] node coordinates in mm (geometry.f90:34-36) */
    const int32_t *nod;          /* [4][neles] NOD(I,K), counter-clockwise              */
    const int32_t *nc;           /* [4][neles] NCk(I): cell id, ghost id, or -1 (then
                                    resolved through bc.in's PARENT/SIDE)               */
    const int32_t *parent, *ghost, *type, *side; /* [nghosts] bc.in columns          */
} tdt_mesh;

/* flow.in fields (turbnstwod.f90:20-23) and the inputs the reference reads but
 * never assigns: BC type 7 PINF, AMINF, DELTABL (boundary.f90:312-317) and BC
 * type 6 TFLAME, ANCOMB, ACOMB, RHOPROP (flame temperature, burn-rate exponent
 * and coefficient, propellant density; boundary.f90:138-153). */
typedef struct tdt_params {
    double gama, amol, cfl, prandtl, prandtlt, p0, t0, pamb, amachinit, rtred;
    double pinf, aminf, deltabl;
    double tflame, ancomb, acomb, rhoprop;
    int32_t local, npseudo, iaxis, reserved;
} tdt_params;

/* ---- lifecycle ---------------------------------------------------------------- */
/* Validates and renumbers the mesh, uploads it, computes the geometry on the GPU
 * (geometry.f90:55-149) and the thermodynamic constants (turbnstwod.f90:37-42).
 * precision: TDT_FP64 or TDT_FP32_MIXED; ordering: TDT_ORDER_*.            */
int tdt_create(tdt_ctx **ctx, const tdt_mesh *mesh, const tdt_params *params,
               int32_t precision, int32_t ordering);
int tdt_destroy(tdt_ctx *ctx);
const char *tdt_last_error(const tdt_ctx *ctx);
const char *tdt_version(void);
/* info[0..7] = neles, nghosts, nods, nmax (largest cell/ghost id), bandwidth of the
 * internal order, precision, ordering actually used, device index */
int tdt_info(const tdt_ctx *ctx, int64_t *info);

/* ---- state -------------------------------------------------------------------- */
/* Uniform start state; returns RINIT (turbnstwod.f90:53-71); COLD <- C. */
int tdt_cold_start(tdt_ctx *ctx, double *rinit);
/* Restart state c = [5][neles] (C1..C5) and RINIT (turbnstwod.f90:73-90); COLD <- C. */
int tdt_set_state(tdt_ctx *ctx, const double *c, double rinit);
int tdt_get_state(tdt_ctx *ctx, double *c);

/* ---- routine-level entry points (FP64; same semantics as the reference) -------- */
int tdt_timestep(tdt_ctx *ctx);            /* turbnstwod.f90:97-114               */
int tdt_dissipation(tdt_ctx *ctx);         /* DISSIPATION   solver_state.f90:70   */
int tdt_primitives(tdt_ctx *ctx);          /* COMPUTE_PRIMITIVE_VARS turbnstwod.f90:204 */
int tdt_apply_boundary(tdt_ctx *ctx);      /* APPLY_BOUNDARY boundary.f90:24      */
int tdt_viscous(tdt_ctx *ctx);             /* COMPUTE_VISCOUS viscous.f90:29      */
int tdt_source_terms(tdt_ctx *ctx);        /* COMPUTE_SOURCE_TERMS source_terms.f90:24 */
int tdt_residual(tdt_ctx *ctx);            /* COMPUTE_RESIDUAL residual.f90:23    */
int tdt_update(tdt_ctx *ctx, double alpha); /* UPDATE_SOLUTION(ALPHA) residual.f90:67 */
/* Per-equation L2 = sqrt(sum (C-COLD)^2 / NELES) and Linf = max |C-COLD|, then
 * COLD <- C (turbnstwod.f90:147-168).  l2, linf: [5] or NULL.  Returns
 * TDT_ERR_NONFINITE if the state holds NaN/Inf (the reference masks it). */
int tdt_monitor(tdt_ctx *ctx, double *l2, double *linf);

/* POSTPROCESS (solver_state.f90:230-283): from the P, U, V, T arrays as
 * COMPUTE_PRIMITIVE_VARS and APPLY_BOUNDARY leave them, the node averages over
 * the cells around each node and the ghosts on its faces, in the reference's
 * summation order, and the tec_plot.dat columns derived from them.
 * nodal: [6][nods] = U, V, P [bar], RHO, MACH, T, in grid.in node order. */
int tdt_postprocess(tdt_ctx *ctx, double *nodal);

/* ---- fused hot path ------------------------------------------------------------- */
/* niter full RK4 iterations (turbnstwod.f90:95-168) on the device: one pre-pass
 * (time step, dissipation sensors, nodal averages) and one fused kernel per RK
 * stage, launched as a HIP graph.  l2, linf: [niter][5] or NULL.
 * first_nonfinite (may be NULL) receives the first iteration (1-based) whose
 * state held NaN/Inf, or 0; the call then returns TDT_ERR_NONFINITE.
 * Leaves C and COLD equal, as at the end of a reference iteration. */
int tdt_iterate(tdt_ctx *ctx, int32_t niter, double *l2, double *linf,
                int32_t *first_nonfinite);

/* ---- residual and its adjoint ------------------------------------------------------
 * No counterpart in the reference (which has no derivative code): the residual
 * operator of one RK stage as a function, and its adjoint, for gradients.
 * R(Q) = R1..R5 of COMPUTE_RESIDUAL (residual.f90:23-62) after
 * COMPUTE_PRIMITIVE_VARS, APPLY_BOUNDARY and COMPUTE_SOURCE_TERMS on the state
 * c = C1..C5 ([5][neles], grid.in cell order), with the context's geometry and
 * flow constants; the context's own state is not touched.  tdt_residual_eval
 * gives r = R(c) ([5][neles]; the same bits as the routine entry points).
 * tdt_residual_vjp gives gq = v^T dR/dQ ([5][neles]) and gparam = v^T dR/d(P0,
 * T0, AMACHINIT) ([3] or NULL; AMACHINIT enters R only through the initial
 * state, so its entry is 0) and gxy = v^T dR/d(X, Y) ([2][nods] or NULL, per
 * grid.in node and millimetre, through the geometry of geometry.f90:55-149).
 * BC type 6 grids and batch contexts: TDT_ERR_UNSUPPORTED. */
int tdt_residual_eval(tdt_ctx *ctx, const double *c, double *r);
int tdt_residual_vjp(tdt_ctx *ctx, const double *c, const double *v, double *gq,
                     double *gparam, double *gxy);
/* New flow.in values for the context (same IAXIS): constants rebuilt, RINIT
 * kept; the next tdt_iterate recaptures its graph. */
int tdt_set_params(tdt_ctx *ctx, const tdt_params *params);
/* New node coordinates (grid.in order, millimetres; same connectivity): the
 * geometry (geometry.f90:55-149) is recomputed. */
int tdt_set_nodes(tdt_ctx *ctx, const double *x_mm, const double *y_mm);

/* ---- adjoint of whole iterations ---------------------------------------------------
 * No counterpart in the reference: the adjoint of the RK4 iteration of
 * turbnstwod.f90:95-145 (time step, DISSIPATION, the four stages with
 * COMPUTE_VISCOUS at stage 1, UPDATE_SOLUTION and its nu~ floor), FP64, with the
 * routine entry points' arithmetic.  Needs LOCAL = 1; BC type 6 grids and batch
 * contexts: TDT_ERR_UNSUPPORTED.  States, adjoints: [5][neles], grid.in order.
 *
 * tdt_operators_eval: out = [10][neles] = D241..D245, VISU, VISV, VIST, VISR, DT
 * of the state c as the adjoint evaluates them (the same bits as the routine
 * entry points after tdt_timestep, tdt_dissipation, tdt_primitives,
 * tdt_apply_boundary and tdt_viscous on c). */
int tdt_operators_eval(tdt_ctx *ctx, const double *c, double *out);
/* Adjoint of those operators at the state c (with RINIT = rinit): for weights
 * w = [7][neles] on E = (D241, D242 + VISU, D243 + VISV, D244 + VIST,
 * D245 + VISR), DT and VOL, g = w^T d(E, DT)/dQ ([5][neles]); gparam ([3] or
 * NULL) = w^T dE/d(P0, T0, RINIT); gxy ([2][nods] or NULL) = w^T d(E, DT,
 * VOL)/d(X, Y) per grid.in millimetre. */
int tdt_operators_vjp(tdt_ctx *ctx, const double *c, double rinit, const double *w, double *g,
                      double *gparam, double *gxy);
/* One iteration Q^n = c (with RINIT = rinit) -> Q^{n+1}: lam = mu^T dQ^{n+1}/dQ^n;
 * gparam ([3] or NULL) = mu^T dQ^{n+1}/d(P0, T0, RINIT); gxy ([2][nods] or NULL)
 * = mu^T dQ^{n+1}/d(X, Y) per grid.in millimetre.  The context's C is kept (COLD
 * is set to it); its routine-level arrays are overwritten. */
int tdt_iteration_vjp(tdt_ctx *ctx, const double *c, double rinit, const double *mu,
                      double *lam, double *gparam, double *gxy);
/* Fixed-point adjoint at a steady state c (with RINIT = rinit; c = Phi(c), Phi one
 * iteration): lam = (I - Phi_Q^T)^-1 dj_dc by the iteration lam <- Phi_Q^T lam +
 * dj_dc from lam = dj_dc, the RK stages recomputed once at c and their blocks
 * kept, until the estimated remaining error relative to the largest |lam| is
 * <= tol or max_iter.  Every 25 iterations the one-step change d (largest
 * |lam_{k+1} - lam_k| over the largest |lam|) is measured; the error estimate is
 * d rho / (1 - rho), with the contraction rate rho from two successive checks
 * (rho^25 = d_k / d_{k-25}).  Then gparam ([3] or NULL) = lam^T
 * dPhi/d(P0, T0, RINIT) and gxy ([2][nods] or NULL) = lam^T dPhi/d(X, Y) per
 * grid.in millimetre, so dJ/dp = dJ/dp|explicit + gparam for the steady state's
 * J.  iters, change (may be NULL): iterations done, the last error estimate (the
 * last one-step change if no rate could be estimated); a value above tol means
 * max_iter came first.  TDT_ERR_NONFINITE if the
 * iteration diverges.  The context's C is kept (COLD is set to it). */
int tdt_adjoint_fixed_point(tdt_ctx *ctx, const double *c, double rinit, const double *dj_dc,
                            int32_t max_iter, double tol, double *lam, double *gparam,
                            double *gxy, int32_t *iters, double *change);
/* niter iterations from c0 with RINIT = rinit (c0 NULL: from the cold start of
 * turbnstwod.f90:53-71, rinit ignored), as tdt_iterate computes them; cn = Q^N.
 * States are kept every ~sqrt(niter) iterations for tdt_adjoint_reverse.  The
 * context's state ends at Q^N. */
int tdt_adjoint_forward(tdt_ctx *ctx, int32_t niter, const double *c0, double rinit,
                        double *cn);
/* Reverse sweep of the last tdt_adjoint_forward, given dj_dcn = dJ/dQ^N:
 * lam0 = dJ/dQ^0; gparam ([4] or NULL) = dJ/d(P0, T0, AMACHINIT, RINIT) -- after
 * a cold start the first three include the path through Q^0 and RINIT (which
 * the cold start computes from them) and the fourth is the part through RINIT
 * alone; from a given c0 they hold Q^0 and RINIT fixed (AMACHINIT: 0) --; gxy
 * ([2][nods] or NULL) = dJ/d(X, Y) per grid.in millimetre.  The stage states of
 * each iteration are recomputed from the kept states, not stored.  tdt_set_params
 * and tdt_set_nodes discard the kept run (TDT_ERR_UNSUPPORTED until the next
 * tdt_adjoint_forward). */
int tdt_adjoint_reverse(tdt_ctx *ctx, const double *dj_dcn, double *lam0, double *gparam,
                        double *gxy);

/* ---- batched cases ----------------------------------------------------------------
 * No counterpart in the reference, which runs one case per process: for parameter
 * sweeps, ncases independent grids advanced together.  One fused iteration (the
 * kernels of tdt_iterate) runs over the concatenated cells of all cases; every
 * cell block holds cells of one case and runs with that case's flow constants,
 * so many small grids fill the GPU together.  Each case is renumbered as
 * tdt_create(ordering) renumbers it alone, and its blocks are those of its own
 * context: states and monitor norms are those of a single-case context, bit for
 * bit in FP64.  The cases share LOCAL (= 1), IAXIS, GAMA, AMOL, CFL, PRANDTL,
 * PRANDTLT and NPSEUDO; P0, T0, AMACHINIT, PAMB, RTRED and the type-7 inputs are
 * per case; BC type 6 is not available (TDT_ERR_UNSUPPORTED).  The routine-level
 * entry points, tdt_cold_start, tdt_set_state, tdt_get_state, tdt_iterate and
 * tdt_postprocess refuse a batch context (TDT_ERR_UNSUPPORTED). */
int tdt_batch_create(tdt_ctx **ctx, int32_t ncases, const tdt_mesh *meshes,
                     const tdt_params *params, int32_t precision, int32_t ordering);
/* Cold start of every case (turbnstwod.f90:53-71 with its own flow.in values);
 * rinit: [ncases] or NULL. */
int tdt_batch_cold_start(tdt_ctx *ctx, double *rinit);
/* State of case icase: c = [5][neles of that case], its grid.in cell order. */
int tdt_batch_set_state(tdt_ctx *ctx, int32_t icase, const double *c, double rinit);
int tdt_batch_get_state(tdt_ctx *ctx, int32_t icase, double *c);
/* niter iterations of every case; l2, linf: [niter][ncases][5] or NULL;
 * first_nonfinite: the first iteration in which any case held NaN/Inf, or 0. */
int tdt_batch_iterate(tdt_ctx *ctx, int32_t niter, double *l2, double *linf,
                      int32_t *first_nonfinite);

/* ---- arrays (reference names, original ids) --------------------------------------
 * Cell+ghost arrays have nmax entries (index = id-1; ids that are neither cell nor
 * ghost read 0): C1..C5 U V T P DT F1..F5 G1..G5 D2Q1..D2Q5 ANU ANURT XCC YCC VOL.
 * Per-cell arrays have neles entries: COLD1..5 R1..R5 D241..D245 VISU VISV VIST
 * VISR SOURCE3 SOURCE5 DL SC1X SC1Y .. SC4X SC4Y.  Node arrays have nods entries:
 * UN VN TN RHON RN.  Scalars (len 1): DTMIN RTOL RINIT.                           */
int tdt_get_array(tdt_ctx *ctx, const char *name, double *out, int32_t len);
int tdt_set_array(tdt_ctx *ctx, const char *name, const double *in, int32_t len);

/* ---- measurement -------------------------------------------------------------- */
/* Average device time (ms) per launch of each fused kernel over the last
 * tdt_iterate call, measured with HIP events on the library's stream when
 * enabled (kernels are then launched eagerly instead of as a graph; 0
 * disables, the default).  times: [6] = pre-pass, dissipation+viscous,
 * stage 1, stage 2, stage 3, stage 4. */
int tdt_set_profiling(tdt_ctx *ctx, int32_t enable);
int tdt_kernel_times(tdt_ctx *ctx, double *times_ms);
/* Kernel set the fused path uses, or -1 before the first tdt_iterate: a choice
 * per kernel slot, as bits.  Every choice computes the same values (FP64: the
 * same bits, except that stages 2-4 in quad form sum the L2 norms in another
 * order; FP32-mixed: the same to FP32 rounding).  The first tdt_iterate call
 * times the uniform sets kernel by kernel on the current state (restoring it
 * after each trial), composes the fastest choice per slot and keeps whichever
 * of that and the uniform sets runs an iteration fastest, unless the
 * environment variable TDT_KERNELS (merged, split, quad, face or a set number), read
 * by the first tdt_iterate, pins one. */
#define TDT_KERNELS_PRE_QUAD 1        /* pre-pass: four lanes per cell (else one)     */
#define TDT_KERNELS_STAGE1_SPLIT 2    /* dissipation + viscous in a kernel of their own */
#define TDT_KERNELS_STAGE1_QUAD 4     /* stage 1 (with D and VIS): four lanes per cell */
#define TDT_KERNELS_STAGE1_FACE 6     /* FP32-mixed single cases: each viscous face
                                         evaluated once (k_vface, the other cell applies
                                         it negated), D and the gathered VIS in stage 1;
                                         equal to the per-cell form to FP32 rounding   */
#define TDT_KERNELS_STAGES_QUAD 8     /* stages 2-4: four lanes per cell               */
#define TDT_KERNELS_PRELOAD 16        /* FP64 one-lane stage kernels (split stage 1,
                                         stages 2-4) request all neighbour states up
                                         front: one memory round trip instead of four,
                                         at one wavefront per SIMD                     */
#define TDT_KERNELS_NODE_LDS 32       /* FP64 split stage 1, single case, LOCAL = 1:
                                         each wavefront of the dissipation / viscous
                                         kernel copies its cells' node records into
                                         LDS (direct global-to-LDS loads) instead of
                                         gathering them face by face                   */
#define TDT_KERNELS_MERGED 0          /* uniform sets: one lane per cell, D/VIS in stage 1 */
#define TDT_KERNELS_SPLIT 2
#define TDT_KERNELS_QUAD 13
#define TDT_KERNELS_FACE 6
int tdt_kernel_set(tdt_ctx *ctx, int32_t *set);
int tdt_synchronize(tdt_ctx *ctx);

/* ---- monitor without allocation ---------------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc), the monitor of tdt_monitor (turbnstwod.f90:147-168) on their own
 * device arrays, with no context.  c: device [5][ldc] = C1..C5 of n cells
 * (component stride ldc >= n); cold: device [5][n] = COLD1..COLD5; out: device
 * [11] = the five L2 = sqrt(sum (C-COLD)^2 / n), the five Linf = max |C-COLD|
 * and a non-finite flag (1.0 if C holds NaN/Inf, else 0.0); save_old != 0 then
 * sets COLD <- C.  Over the same cells in the same order the values have the
 * same bits as tdt_monitor's.  Asynchronous on stream (a hipStream_t; NULL =
 * the null stream): out is valid once the stream has reached it.  Allocates
 * nothing: the per-block partials live in a static device array of
 * TDT_MONITOR_NOALLOC_MAX_CELLS / 256 records (5.8 MB), one per GPU, held until
 * the process ends; n above TDT_MONITOR_NOALLOC_MAX_CELLS gives TDT_ERR_ARG.
 * All calls on one GPU share that array: calls that overlap on one GPU race.
 * Errors: tdt_last_error(NULL). */
#define TDT_MONITOR_NOALLOC_MAX_CELLS (1 << 24)
int tdt_monitor_noalloc(const double *c, int64_t ldc, double *cold, int32_t n, int32_t save_old,
                        double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* TDT_H */
