/*
 * tdt_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C, FP64) of the reference solver's per-iteration
 * residual and RK4 update path.  Only the tests, the smoke check and the CPU
 * baseline leg of the benchmark load it, and only as the checker; the product
 * (lib/libtdt.so) never links or calls it.
 *
 * Arrays keep the reference layout: 1-based ids, cells 1..NELES, ghosts at
 * the ids bc.in gives them, nodes 1..NODS.  Index 0 of every array is unused.
 */
#ifndef TDT_ORACLE_H
#define TDT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

/* dparams: gama amol cfl prandtl prandtlt p0 t0 pamb amachinit rtred pinf aminf deltabl
 *          tflame ancomb acomb rhoprop
 * iparams: local npseudo iaxis
 * nod, nc: column-major [4][neles], 1-based ids (NOD(I,K) = nod[(K-1)*neles + I-1]). */
orc *orc_create(int32_t nods, int32_t neles, int32_t nghosts, const double *x_mm,
                const double *y_mm, const int32_t *nod, const int32_t *nc,
                const int32_t *parent, const int32_t *ghost, const int32_t *type,
                const int32_t *side, const double *dparams, const int32_t *iparams);
void orc_destroy(orc *o);
const char *orc_error(void);

void orc_cold_start(orc *o);                               /* turbnstwod.f90:53-71  */
void orc_set_state(orc *o, const double *c, double rinit); /* turbnstwod.f90:73-90  */
void orc_timestep(orc *o);                                 /* turbnstwod.f90:97-114 */
void orc_dissipation(orc *o);                              /* solver_state.f90:70-225 */
void orc_primitives(orc *o);                               /* turbnstwod.f90:204-227 */
void orc_apply_boundary(orc *o);                           /* boundary.f90:24-48    */
void orc_viscous(orc *o);                                  /* viscous.f90:29-217    */
void orc_source_terms(orc *o);                             /* source_terms.f90:24-104 */
void orc_residual(orc *o);                                 /* residual.f90:23-62    */
void orc_update(orc *o, double alpha);                     /* residual.f90:67-83    */
/* L2 = sqrt(sum (C-COLD)^2 / NELES) and Linf = max |C-COLD| per equation,
 * then COLD <- C (turbnstwod.f90:147-168). */
void orc_monitor(orc *o, double *l2, double *linf);
/* niter full iterations; l2/linf may be NULL, else [niter][5]. */
void orc_iterate(orc *o, int32_t niter, double *l2, double *linf);

/* Pointer to element 1 of a named internal array; *len = its length.
 * Names: C1..C5 COLD1..5 U V T P DT F1..5 G1..5 R1..5 D2Q1..5 D241..5 ANU ANURT
 * VISU VISV VIST VISR SOURCE3 SOURCE5 XCC YCC VOL DL SC1X..SC4Y (nmax long or
 * neles for per-cell-only arrays), UN VN TN RN RHON (nods long). */
double *orc_array(orc *o, const char *name, int32_t *len);
double orc_scalar(orc *o, const char *name);  /* RTOL RINIT DTMIN RR CV CP S0 TINF */

#ifdef __cplusplus
}
#endif
#endif
