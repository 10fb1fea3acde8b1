/*
 * tdt_oracle.c -- TEST INFRASTRUCTURE ONLY (see tdt_oracle.h).
 *
 * A plain-C, FP64 restatement of the reference solver's hot path.  Every
 * expression keeps the reference's evaluation order (Fortran evaluates
 * left to right within a precedence level and honours parentheses), and the
 * file is compiled with -ffp-contract=off, so on x86-64 with glibc libm the
 * results agree bit for bit with the reference built by amdflang -O3 (the
 * golden fixtures under tests/golden check exactly that).
 *
 * Each function cites the reference lines it restates.
 */
#include "tdt_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* constants.f90:12-45 */
#define RUNIV_VAL 8314.0
#define CMU_VAL 0.09
#define ALPHA_TURB 0.00052
#define BETA_TURB 0.2
#define CONST1_VAL 21.479
#define CONST2_VAL 21.394
#define CONST3_VAL 1.5
#define SUTH_COEF 0.00000011848
#define C02_VAL 0.25
#define C04_VAL 0.00390625
#define CFLTURB_VAL 0.5
#define QUARTER 0.25
#define HALF 0.5
#define MM_TO_M 0.001
#define YAXIS_THRESH 0.001
#define AMU_THRESH 0.001
static const double DIV23_VAL = 2.0 / 3.0;
static const double BL_EXPONENT = 1.0 / 7.0;
#define CONVKSC_VAL (1.0 / 98066.5)   /* constants.f90:55 */
#define GTH_VAL 35.0                  /* constants.f90:56 */

static char g_err[256];
const char *orc_error(void) { return g_err; }

struct orc {
    int nods, neles, nghosts, nmax;
    double *x, *y;
    int *nod[4];
    int *nc[4];
    int *parent, *ghost, *type, *side;
    double gama, amol, cfl, prandtl, prandtlt, p0, t0, pamb, amachinit, rtred, pinf, aminf,
        deltabl, tflame, ancomb, acomb, rhoprop;
    int local, npseudo, iaxis;
    double rr, cv, cp, rho0, s0, tinf, rinit, rtol, dtmin;
    double *xcc, *ycc, *vol, *dl, *sc[8];
    double *c[5], *cold[5], *u, *v, *t, *p, *dt, *f[5], *g[5], *r[5], *d2q[5], *d24[5];
    double *anu, *anurt, *vis[4], *src3, *src5;
    double *un, *vn, *tn, *rn, *rhon;
    int *itest;
};

static double *dalloc(int n) { return (double *)calloc((size_t)n + 1, sizeof(double)); }
static int *ialloc(int n) { return (int *)calloc((size_t)n + 1, sizeof(int)); }

/* integer power by binary exponentiation, as Fortran real**integer */
static double ipow(double a, int b) {
    int neg = b < 0;
    unsigned ub = neg ? (unsigned)(-(long)b) : (unsigned)b;
    double r = 1.0;
    while (ub) {
        if (ub & 1u) r *= a;
        ub >>= 1;
        if (ub) a *= a;
    }
    return neg ? 1.0 / r : r;
}

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* geometry.f90:22-151 (file reading is done by the caller) */
static void geometry(orc *o) {
    for (int i = 1; i <= o->neles; ++i) {
        int n1 = o->nod[0][i], n2 = o->nod[1][i], n3 = o->nod[2][i], n4 = o->nod[3][i];
        o->itest[n1]++; o->itest[n2]++; o->itest[n3]++; o->itest[n4]++;
        double x1 = o->x[n1], x2 = o->x[n2], x3 = o->x[n3], x4 = o->x[n4];
        double y1 = o->y[n1], y2 = o->y[n2], y3 = o->y[n3], y4 = o->y[n4];
        o->xcc[i] = (x1 + x2 + x3 + x4) * 0.25;
        o->ycc[i] = (y1 + y2 + y3 + y4) * 0.25;
        double dl12 = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
        double dl23 = sqrt((x2 - x3) * (x2 - x3) + (y2 - y3) * (y2 - y3));
        double dl34 = sqrt((x3 - x4) * (x3 - x4) + (y3 - y4) * (y3 - y4));
        double dl41 = sqrt((x4 - x1) * (x4 - x1) + (y4 - y1) * (y4 - y1));
        o->dl[i] = dmin(dmin(dmin(dl12, dl23), dl34), dl41);
        if (o->iaxis == 1) {
            o->sc[0][i] = (y2 - y1) * (y1 + y2) * HALF;
            o->sc[1][i] = -((x2 - x1) * (y1 + y2) * HALF);
            o->sc[2][i] = (y3 - y2) * (y2 + y3) * HALF;
            o->sc[3][i] = -((x3 - x2) * (y2 + y3) * HALF);
            o->sc[4][i] = (y4 - y3) * (y3 + y4) * HALF;
            o->sc[5][i] = -((x4 - x3) * (y3 + y4) * HALF);
            o->sc[6][i] = (y1 - y4) * (y4 + y1) * HALF;
            o->sc[7][i] = -((x1 - x4) * (y4 + y1) * HALF);
            o->vol[i] = HALF * ((x3 - x1) * (y4 - y2) - (x4 - x2) * (y3 - y1)) * o->ycc[i];
        } else {
            o->sc[0][i] = (y2 - y1);
            o->sc[1][i] = -(x2 - x1);
            o->sc[2][i] = (y3 - y2);
            o->sc[3][i] = -(x3 - x2);
            o->sc[4][i] = (y4 - y3);
            o->sc[5][i] = -(x4 - x3);
            o->sc[6][i] = (y1 - y4);
            o->sc[7][i] = -(x1 - x4);
            o->vol[i] = HALF * ((x3 - x1) * (y4 - y2) - (x4 - x2) * (y3 - y1));
        }
        if (o->vol[i] <= 0.0) o->vol[i] = -o->vol[i];
    }
    for (int gi = 1; gi <= o->nghosts; ++gi) {
        int np = o->parent[gi], ng = o->ghost[gi], ns = o->side[gi];
        int a = o->nod[(ns - 1) & 3][np], b = o->nod[ns & 3][np];
        double xf = HALF * (o->x[a] + o->x[b]);
        double yf = HALF * (o->y[a] + o->y[b]);
        o->xcc[ng] = 2.0 * xf - o->xcc[np];
        o->ycc[ng] = 2.0 * yf - o->ycc[np];
        o->itest[a]++; o->itest[b]++;
        o->vol[ng] = o->vol[np];
        for (int k = 0; k < 4; ++k) o->nc[k][ng] = ng;
    }
}

orc *orc_create(int32_t nods, int32_t neles, int32_t nghosts, const double *x_mm,
                const double *y_mm, const int32_t *nod, const int32_t *nc,
                const int32_t *parent, const int32_t *ghost, const int32_t *type,
                const int32_t *side, const double *dp, const int32_t *ip) {
    g_err[0] = 0;
    int nmax = neles;
    for (int g = 0; g < nghosts; ++g) {
        if (ghost[g] <= neles) {
            snprintf(g_err, sizeof g_err, "ghost id %d collides with cell ids", ghost[g]);
            return NULL;
        }
        if (type[g] < 2 || type[g] > 7) {
            snprintf(g_err, sizeof g_err, "boundary type %d not supported", type[g]);
            return NULL;
        }
        if (ghost[g] > nmax) nmax = ghost[g];
    }
    orc *o = (orc *)calloc(1, sizeof(orc));
    o->nods = nods; o->neles = neles; o->nghosts = nghosts; o->nmax = nmax;
    o->x = dalloc(nods); o->y = dalloc(nods);
    for (int i = 0; i < nods; ++i) {
        o->x[i + 1] = x_mm[i] * MM_TO_M;
        o->y[i + 1] = y_mm[i] * MM_TO_M;
    }
    for (int k = 0; k < 4; ++k) {
        o->nod[k] = ialloc(neles);
        o->nc[k] = ialloc(nmax);
        for (int i = 0; i < neles; ++i) {
            o->nod[k][i + 1] = nod[(size_t)k * neles + i];
            o->nc[k][i + 1] = nc[(size_t)k * neles + i];
        }
    }
    o->parent = ialloc(nghosts); o->ghost = ialloc(nghosts);
    o->type = ialloc(nghosts); o->side = ialloc(nghosts);
    for (int g = 0; g < nghosts; ++g) {
        o->parent[g + 1] = parent[g]; o->ghost[g + 1] = ghost[g];
        o->type[g + 1] = type[g]; o->side[g + 1] = side[g];
    }
    o->gama = dp[0]; o->amol = dp[1]; o->cfl = dp[2]; o->prandtl = dp[3]; o->prandtlt = dp[4];
    o->p0 = dp[5]; o->t0 = dp[6]; o->pamb = dp[7]; o->amachinit = dp[8]; o->rtred = dp[9];
    o->pinf = dp[10]; o->aminf = dp[11]; o->deltabl = dp[12];
    o->tflame = dp[13]; o->ancomb = dp[14]; o->acomb = dp[15]; o->rhoprop = dp[16];
    o->local = ip[0]; o->npseudo = ip[1]; o->iaxis = ip[2];

    o->xcc = dalloc(nmax); o->ycc = dalloc(nmax); o->vol = dalloc(nmax); o->dl = dalloc(nmax);
    for (int k = 0; k < 8; ++k) o->sc[k] = dalloc(nmax);
    for (int m = 0; m < 5; ++m) {
        o->c[m] = dalloc(nmax); o->cold[m] = dalloc(nmax); o->f[m] = dalloc(nmax);
        o->g[m] = dalloc(nmax); o->r[m] = dalloc(nmax); o->d2q[m] = dalloc(nmax);
        o->d24[m] = dalloc(nmax);
    }
    o->u = dalloc(nmax); o->v = dalloc(nmax); o->t = dalloc(nmax); o->p = dalloc(nmax);
    o->dt = dalloc(nmax); o->anu = dalloc(nmax); o->anurt = dalloc(nmax);
    for (int k = 0; k < 4; ++k) o->vis[k] = dalloc(nmax);
    o->src3 = dalloc(nmax); o->src5 = dalloc(nmax);
    o->un = dalloc(nods); o->vn = dalloc(nods); o->tn = dalloc(nods); o->rn = dalloc(nods);
    o->rhon = dalloc(nods);
    o->itest = ialloc(nmax > nods ? nmax : nods);

    geometry(o);
    /* thermodynamics, turbnstwod.f90:37-42 */
    o->rr = RUNIV_VAL / o->amol;
    o->cv = o->rr / (o->gama - 1.0);
    o->cp = o->cv * o->gama;
    o->rho0 = o->p0 / (o->rr * o->t0);
    o->s0 = o->p0 / pow(o->rho0, o->gama);
    o->tinf = o->t0 / (1.0 + 0.5 * (o->gama - 1.0) * o->aminf * o->aminf);
    return o;
}

void orc_destroy(orc *o) {
    if (!o) return;
    free(o->x); free(o->y);
    for (int k = 0; k < 4; ++k) { free(o->nod[k]); free(o->nc[k]); free(o->vis[k]); }
    free(o->parent); free(o->ghost); free(o->type); free(o->side);
    free(o->xcc); free(o->ycc); free(o->vol); free(o->dl);
    for (int k = 0; k < 8; ++k) free(o->sc[k]);
    for (int m = 0; m < 5; ++m) {
        free(o->c[m]); free(o->cold[m]); free(o->f[m]); free(o->g[m]); free(o->r[m]);
        free(o->d2q[m]); free(o->d24[m]);
    }
    free(o->u); free(o->v); free(o->t); free(o->p); free(o->dt); free(o->anu); free(o->anurt);
    free(o->src3); free(o->src5);
    free(o->un); free(o->vn); free(o->tn); free(o->rn); free(o->rhon); free(o->itest);
    free(o);
}

static void save_old(orc *o) {
    for (int m = 0; m < 5; ++m)
        for (int i = 1; i <= o->neles; ++i) o->cold[m][i] = o->c[m][i];
}

/* turbnstwod.f90:53-71, then COLD <- C (:87-90) */
void orc_cold_start(orc *o) {
    double g = o->gama;
    double dynratio = 1.0 + (g - 1.0) * 0.5 * (o->amachinit * o->amachinit);
    double tinit = o->t0 / dynratio;
    double pinit = o->p0 / pow(dynratio, g / (g - 1.0));
    double uinit = o->amachinit * sqrt(g * o->rr * tinit);
    double vinit = 0.0;
    double einit = o->cv * tinit + 0.5 * (uinit * uinit + vinit * vinit);
    double rhoinit = pinit / (o->rr * tinit);
    double amuinit = (0.00000011848) * sqrt(o->amol) * pow(tinit, 0.6);
    o->rinit = amuinit / rhoinit;
    o->rtol = o->rinit * o->rtred;
    for (int i = 1; i <= o->neles; ++i) {
        o->c[0][i] = rhoinit;
        o->c[1][i] = rhoinit * uinit;
        o->c[2][i] = rhoinit * vinit;
        o->c[3][i] = rhoinit * einit;
        o->c[4][i] = rhoinit * o->rinit;
    }
    save_old(o);
}

/* turbnstwod.f90:73-90 (restart: state and RINIT supplied by the caller) */
void orc_set_state(orc *o, const double *c, double rinit) {
    for (int m = 0; m < 5; ++m)
        for (int i = 0; i < o->neles; ++i) o->c[m][i + 1] = c[(size_t)m * o->neles + i];
    o->rinit = rinit;
    o->rtol = rinit * o->rtred;
    save_old(o);
}

/* turbnstwod.f90:97-114 */
void orc_timestep(orc *o) {
    double dtmin = 100000.0;
    for (int i = 1; i <= o->neles; ++i) {
        double rho = o->c[0][i];
        o->u[i] = o->c[1][i] / rho;
        o->v[i] = o->c[2][i] / rho;
        double vel = sqrt(o->u[i] * o->u[i] + o->v[i] * o->v[i]);
        double e = o->c[3][i] / rho;
        o->t[i] = (e - 0.5 * (o->u[i] * o->u[i] + o->v[i] * o->v[i])) / o->cv;
        o->p[i] = rho * o->rr * o->t[i];
        o->dt[i] = o->cfl * o->dl[i] / (fabs(vel) + sqrt(o->gama * o->rr * o->t[i]));
        dtmin = dmin(dtmin, o->dt[i]);
    }
    o->dtmin = dtmin;
    if (o->local == 0)
        for (int i = 1; i <= o->neles; ++i) o->dt[i] = dtmin;
}

/* solver_state.f90:70-225 */
void orc_dissipation(orc *o) {
    for (int gi = 1; gi <= o->nghosts; ++gi) {
        int ng = o->ghost[gi], np = o->parent[gi];
        o->p[ng] = o->p[np];
        o->dt[ng] = o->dt[np];
        for (int m = 0; m < 5; ++m) o->c[m][ng] = o->c[m][np];
    }
    for (int i = 1; i <= o->neles; ++i) {
        int n1 = o->nc[0][i], n2 = o->nc[1][i], n3 = o->nc[2][i], n4 = o->nc[3][i];
        double pie = o->p[i], c1i = o->c[0][i], c5i = o->c[4][i];
        const double *P = o->p, *C5 = o->c[4];
        double anun = fabs(P[n1] - pie) + fabs(P[n2] - pie) + fabs(P[n3] - pie) + fabs(P[n4] - pie);
        double anud = fabs(P[n1] + pie) + fabs(P[n2] + pie) + fabs(P[n3] + pie) + fabs(P[n4] + pie);
        o->anu[i] = anun / anud;
        anun = fabs(C5[n1] - c5i) + fabs(C5[n2] - c5i) + fabs(C5[n3] - c5i) + fabs(C5[n4] - c5i);
        anud = fabs(C5[n1] + c5i) + fabs(C5[n2] + c5i) + fabs(C5[n3] + c5i) + fabs(C5[n4] + c5i);
        o->anurt[i] = anun / dmax(anud, 8.0 * c1i * o->rtol);
        for (int m = 0; m < 5; ++m) {
            const double *C = o->c[m];
            o->d2q[m][i] = C[n1] + C[n2] + C[n3] + C[n4] - 4.0 * C[i];
        }
    }
    for (int gi = 1; gi <= o->nghosts; ++gi) {
        int ng = o->ghost[gi], np = o->parent[gi];
        o->anu[ng] = o->anu[np];
        for (int m = 0; m < 5; ++m) o->d2q[m][ng] = o->d2q[m][np];
    }
    for (int i = 1; i <= o->neles; ++i) {
        double volbydt = o->vol[i] / o->dt[i];
        int nb[4] = {o->nc[0][i], o->nc[1][i], o->nc[2][i], o->nc[3][i]};
        for (int m = 0; m < 5; ++m) {
            const double *sens = (m < 4) ? o->anu : o->anurt;
            double si = sens[i];
            const double *C = o->c[m], *Q = o->d2q[m];
            double d2 = 0.0, d4 = 0.0;
            for (int k = 0; k < 4; ++k) {
                int n = nb[k];
                double e2n = C02_VAL * dmax(si, sens[n]);
                double e4n = dmax(0.0, C04_VAL - e2n);
                double d24n = HALF * (volbydt + o->vol[n] / o->dt[n]);
                double d2n = e2n * d24n, d4n = e4n * d24n;
                if (k == 0) {
                    d2 = d2n * (C[n] - C[i]);
                    d4 = d4n * (Q[n] - Q[i]);
                } else {
                    d2 = d2 + d2n * (C[n] - C[i]);
                    d4 = d4 + d4n * (Q[n] - Q[i]);
                }
            }
            o->d24[m][i] = d2 - d4;
        }
    }
}

/* turbnstwod.f90:204-227 */
void orc_primitives(orc *o) {
    for (int i = 1; i <= o->neles; ++i) {
        double rho = o->c[0][i];
        double u = o->c[1][i] / rho, v = o->c[2][i] / rho;
        double e = o->c[3][i] / rho;
        double t = (e - 0.5 * (u * u + v * v)) / o->cv;
        double p = rho * o->rr * t;
        o->u[i] = u; o->v[i] = v; o->t[i] = t; o->p[i] = p;
        o->f[0][i] = o->c[1][i];
        o->f[1][i] = rho * u * u + p;
        o->f[2][i] = rho * u * v;
        o->f[3][i] = (o->c[3][i] + p) * u;
        o->g[0][i] = o->c[2][i];
        o->g[1][i] = o->f[2][i];
        o->g[2][i] = rho * v * v + p;
        o->g[3][i] = (o->c[3][i] + p) * v;
        double rt = o->c[4][i] / rho;
        o->f[4][i] = rho * rt * u;
        o->g[4][i] = rho * rt * v;
    }
}

/* Ghost conservative variables and fluxes 1-4 from ghost U, V, T, P, RHO
 * (the common tail of boundary.f90:62-69, 89-93, 103-107, 321-325). */
static void ghost_fill(orc *o, int ng, double rho) {
    double u = o->u[ng], v = o->v[ng], t = o->t[ng], p = o->p[ng];
    o->c[0][ng] = rho;
    o->c[1][ng] = rho * u;
    o->c[2][ng] = rho * v;
    double e = o->cv * t + 0.5 * (u * u + v * v);
    o->c[3][ng] = rho * e;
    o->f[0][ng] = o->c[1][ng];
    o->f[1][ng] = rho * u * u + p;
    o->f[2][ng] = rho * u * v;
    o->f[3][ng] = (o->c[3][ng] + p) * u;
    o->g[0][ng] = o->c[2][ng];
    o->g[1][ng] = o->f[2][ng];
    o->g[2][ng] = rho * v * v + p;
    o->g[3][ng] = (o->c[3][ng] + p) * v;
}

/* DELY of cell n with the node-pair expression of face `side`
 * (boundary.f90:176-205, 229-237) */
static double dely_face(const orc *o, int n, int side) {
    const double y1 = o->y[o->nod[0][n]], y2 = o->y[o->nod[1][n]];
    const double y3 = o->y[o->nod[2][n]], y4 = o->y[o->nod[3][n]];
    switch (side) {
    case 1: return HALF * ((y1 + y2) - (y3 + y4));
    case 2: return HALF * ((y2 + y3) - (y1 + y4));
    case 3: return HALF * ((y3 + y4) - (y1 + y2));
    default: return HALF * ((y4 + y1) - (y2 + y3));
    }
}

static double gport_term(const orc *o, int n, double dely) {
    return o->iaxis == 1 ? o->c[0][n] * o->u[n] * dely * o->ycc[n] : o->c[0][n] * o->u[n] * dely;
}

/* COMPUTE_GPORT + ACCUMULATE_GPORT_ALONG_BOUNDARY (boundary.f90:169-305): the
 * mass flow through the column of cells from the parent across to the far
 * boundary (the walk ends at the ghost whose neighbour ids point to itself,
 * geometry.f90:145-148), normalised by the hydraulic size DHYD */
static double gport(const orc *o, int np, int ns, double dhyd) {
    static const int walk_nc[5] = {0, 2, 3, 0, 1};   /* NS 1->NC3, 2->NC4, 3->NC1, 4->NC2 */
    static const int walk_side[5] = {0, 3, 4, 1, 2};
    double g = gport_term(o, np, dely_face(o, np, ns));
    const int side = walk_side[ns];
    int n = o->nc[walk_nc[ns]][np];
    g = g + gport_term(o, n, dely_face(o, n, side));
    const int *ncw = o->nc[side - 1];
    n = ncw[n];
    /* SIDE 1 walks NC1 with face 3's expression, 2: NC2 / face 4, 3: NC3 / face 1,
     * 4: NC4 / face 2 */
    const int lside = (side + 1) % 4 + 1;
    while (ncw[n] != n) {
        g = g + gport_term(o, n, dely_face(o, n, lside));
        n = ncw[n];
    }
    return o->iaxis == 1 ? 2.0 * g / (dhyd * dhyd) : g / dhyd;
}

/* boundary.f90:24-122, 128-164, 310-327 */
void orc_apply_boundary(orc *o) {
    for (int gi = 1; gi <= o->nghosts; ++gi) {
        int np = o->parent[gi], ng = o->ghost[gi], nt = o->type[gi], ns = o->side[gi];
        double *U = o->u, *V = o->v, *T = o->t, *P = o->p;
        if (nt == 2) {                                   /* boundary.f90:53-71 */
            U[ng] = fabs(U[np]);
            V[ng] = -V[np];
            T[ng] = o->t0 - 0.5 * (U[ng] * U[ng] + V[ng] * V[ng]) / o->cp;
            double rho = pow(o->s0 / (o->rr * T[ng]), 1.0 / (1.0 - o->gama));
            P[ng] = rho * o->rr * T[ng];
            double amu = (0.00000011848) * sqrt(o->amol) * pow(T[ng], 0.6);
            double rt = amu / rho;
            ghost_fill(o, ng, rho);
            o->c[4][ng] = rho * rt;
            o->f[4][ng] = o->c[4][ng] * U[ng];
            o->g[4][ng] = o->c[4][ng] * V[ng];
        } else if (nt == 3) {                            /* boundary.f90:76-95 */
            int ncopp = o->nc[(ns + 1) & 3][np];         /* side 1->NC3, 2->NC4, 3->NC1, 4->NC2 */
            U[ng] = fabs(2.0 * U[np] - U[ncopp]);
            V[ng] = 2.0 * V[np] - V[ncopp];
            T[ng] = 2.0 * T[np] - T[ncopp];
            P[ng] = 2.0 * P[np] - P[ncopp];
            if (o->pamb >= 0.0) P[ng] = o->pamb;
            double rho = P[ng] / (o->rr * T[ng]);
            ghost_fill(o, ng, rho);
            o->c[4][ng] = o->c[4][np];
            o->f[4][ng] = o->f[4][np];
            o->g[4][ng] = o->g[4][np];
        } else if (nt == 4) {                            /* boundary.f90:100-109 */
            U[ng] = U[np]; V[ng] = -V[np]; T[ng] = T[np];
            double rho = o->c[0][np];
            P[ng] = P[np];
            ghost_fill(o, ng, rho);
            o->c[4][ng] = o->c[4][np];
            o->f[4][ng] = o->f[4][np];
            o->g[4][ng] = -o->g[4][np];
        } else if (nt == 5) {                            /* boundary.f90:114-122 */
            U[ng] = -U[np]; V[ng] = -V[np]; T[ng] = T[np];
            double rho = o->c[0][np];
            P[ng] = P[np];
            o->c[0][ng] = rho;
            o->c[1][ng] = rho * U[ng];
            o->c[2][ng] = rho * V[ng];
            double e = o->cv * T[ng] + 0.5 * (U[ng] * U[ng] + V[ng] * V[ng]);
            o->c[3][ng] = rho * e;
            o->f[0][ng] = -o->f[0][np];
            o->f[1][ng] = 2.0 * P[np] - o->f[1][np];
            o->f[2][ng] = -o->f[2][np];
            o->f[3][ng] = -o->f[3][np];
            o->g[0][ng] = -o->g[0][np];
            o->g[1][ng] = -o->g[1][np];
            o->g[2][ng] = 2.0 * P[np] - o->g[2][np];
            o->g[3][ng] = -o->g[3][np];
            o->c[4][ng] = -o->c[4][np];
            o->f[4][ng] = -o->f[4][np];
            o->g[4][ng] = -o->g[4][np];
        } else if (nt == 6) {                            /* boundary.f90:128-164 */
            const double dhyd = fabs(o->ycc[ng] + o->ycc[np]);
            const double scx = o->sc[2 * (ns - 1)][np], scy = o->sc[2 * (ns - 1) + 1][np];
            const double dnr = sqrt(scx * scx + scy * scy);
            const double enx = -scx / dnr, eny = -scy / dnr;
            const double gp = gport(o, np, ns, dhyd);
            double rho = o->c[0][np];
            const double rdot = o->acomb * pow(P[np] * CONVKSC_VAL, o->ancomb);
            double rdoteb = rdot;
            double vel = o->rhoprop / rho * rdoteb;
            P[ng] = P[np];
            U[ng] = vel * enx + vel * enx - U[np];
            V[ng] = vel * eny + vel * eny - V[np];
            T[ng] = o->tflame - 0.5 * (U[ng] * U[ng] + V[ng] * V[ng]) / o->cp;
            double amu = (0.00000011848) * sqrt(o->amol) * pow(T[ng], 0.6);
            const double gval = (gp / (o->rhoprop * rdot)) *
                                pow(o->rhoprop * rdot * dhyd / amu, -0.125);
            if (gval > GTH_VAL)
                rdoteb = rdoteb * (1.0 + 0.023 * (pow(gval, 0.8) - pow(GTH_VAL, 0.8)));
            vel = o->rhoprop / rho * rdoteb;
            U[ng] = vel * enx + vel * enx - U[np];
            V[ng] = vel * eny + vel * eny - V[np];
            T[ng] = o->tflame - 0.5 * (U[ng] * U[ng] + V[ng] * V[ng]) / o->cp;
            rho = P[ng] / (o->rr * T[ng]);
            amu = (0.00000011848) * sqrt(o->amol) * pow(T[ng], 0.6);
            const double rt = amu / rho;
            ghost_fill(o, ng, rho);
            o->c[4][ng] = rho * rt;
            o->f[4][ng] = o->c[4][ng] * U[ng];
            o->g[4][ng] = o->c[4][ng] * V[ng];
        } else if (nt == 7) {                            /* boundary.f90:310-327 */
            P[ng] = o->pinf; T[ng] = o->tinf; V[ng] = -V[np];
            double rho = P[ng] / (o->rr * T[ng]);
            if (o->ycc[ng] <= o->deltabl)
                U[ng] = 2.0 * o->aminf * sqrt(o->gama * o->rr * o->tinf) *
                        pow(o->ycc[ng] / o->deltabl, BL_EXPONENT) - U[np];
            else
                U[ng] = 2.0 * o->aminf * sqrt(o->gama * o->rr * o->tinf) - U[np];
            double amu = (0.00000011848) * sqrt(o->amol) * pow(T[ng], 0.6);
            double rt = amu / rho;
            ghost_fill(o, ng, rho);
            o->c[4][ng] = rho * rt;
            o->f[4][ng] = o->c[4][ng] * U[ng];
            o->g[4][ng] = o->c[4][ng] * V[ng];
        }
    }
}

/* viscous.f90:130-217 (one face of cell i) */
static void viscous_face(orc *o, int i, int face, int nn1, int nn2, int ncc) {
    double sx = o->sc[2 * (face - 1)][i], sy = o->sc[2 * (face - 1) + 1][i];
    double xb = o->x[nn1], yb = o->y[nn1], xd = o->x[nn2], yd = o->y[nn2];
    double xa = o->xcc[i], ya = o->ycc[i], xc = o->xcc[ncc], yc = o->ycc[ncc];
    double ua = o->u[i], va = o->v[i], ta = o->t[i];
    double ub = o->un[nn1], vb = o->vn[nn1], tb = o->tn[nn1];
    double uc = o->u[ncc], vc = o->v[ncc], tc = o->t[ncc];
    double ud = o->un[nn2], vd = o->vn[nn2], td = o->tn[nn2];
    double rhoa = o->c[0][i], rhob = o->rhon[nn1], rhoc = o->c[0][ncc], rhod = o->rhon[nn2];
    double ra = o->c[4][i] / o->c[0][i], rb = o->rn[nn1], rc = o->c[4][ncc] / o->c[0][ncc],
           rd = o->rn[nn2];

    double dnr = (xa - xc) * (yb - yd) - (xb - xd) * (ya - yc);
    double dux = ((ua - uc) * (yb - yd) - (ub - ud) * (ya - yc)) / dnr;
    double dvx = ((va - vc) * (yb - yd) - (vb - vd) * (ya - yc)) / dnr;
    double dtx = ((ta - tc) * (yb - yd) - (tb - td) * (ya - yc)) / dnr;
    double drx = ((ra - rc) * (yb - yd) - (rb - rd) * (ya - yc)) / dnr;
    double duy = (-((ua - uc) * (xb - xd)) + (ub - ud) * (xa - xc)) / dnr;
    double dvy = (-((va - vc) * (xb - xd)) + (vb - vd) * (xa - xc)) / dnr;
    double dty = (-((ta - tc) * (xb - xd)) + (tb - td) * (xa - xc)) / dnr;
    double dry = (-((ra - rc) * (xb - xd)) + (rb - rd) * (xa - xc)) / dnr;

    double uavg = QUARTER * (ua + ub + uc + ud);
    double vavg = QUARTER * (va + vb + vc + vd);
    double tavg = QUARTER * (ta + tb + tc + td);
    double rhoavg = QUARTER * (rhoa + rhob + rhoc + rhod);
    double ravg = QUARTER * (ra + rb + rc + rd);

    double divvel;
    if (o->iaxis == 1) {
        double yavg = HALF * (yb + yd);
        double vbyy = (yavg < YAXIS_THRESH) ? dvy : vavg / yavg;
        divvel = dux + dvy + vbyy;
    } else {
        divvel = dux + dvy;
    }
    double amu = SUTH_COEF * sqrt(o->amol) * pow(tavg, 0.6);
    double ak = amu * o->cp / o->prandtl;
    double chi = rhoavg * ravg / (amu * CMU_VAL);
    double fmu = tanh(ALPHA_TURB * chi * chi) / tanh(BETA_TURB * chi * chi);
    double amut = rhoavg * ravg * fmu;
    double akt = amut * o->cp / o->prandtlt;
    double amueff = amu + amut;
    double akeff = ak + akt;
    double txx = amueff * (dux + dux - DIV23_VAL * divvel);
    double tyy = amueff * (dvy + dvy - DIV23_VAL * divvel);
    double txy = amueff * (duy + dvx);
    double qxx = uavg * txx + vavg * txy + akeff * dtx;
    double qyy = uavg * txy + vavg * tyy + akeff * dty;
    if (face == 1) {
        o->vis[0][i] = txx * sx + txy * sy;
        o->vis[1][i] = txy * sx + tyy * sy;
        o->vis[2][i] = qxx * sx + qyy * sy;
        o->vis[3][i] = amueff * drx * sx + amueff * dry * sy;
    } else {
        o->vis[0][i] = o->vis[0][i] + txx * sx + txy * sy;
        o->vis[1][i] = o->vis[1][i] + txy * sx + tyy * sy;
        o->vis[2][i] = o->vis[2][i] + qxx * sx + qyy * sy;
        o->vis[3][i] = o->vis[3][i] + amueff * drx * sx + amueff * dry * sy;
    }
}

/* viscous.f90:29-124 */
void orc_viscous(orc *o) {
    int nn = o->nods;
    for (int n = 1; n <= nn; ++n) {
        o->un[n] = o->vn[n] = o->tn[n] = o->rn[n] = o->rhon[n] = 0.0;
        o->itest[n] = 0;
    }
    for (int i = 1; i <= o->neles; ++i) {                /* viscous.f90:77-93 */
        double ui = o->u[i], vi = o->v[i], ti = o->t[i], rhoi = o->c[0][i];
        double ri = o->c[4][i] / rhoi;
        for (int k = 0; k < 4; ++k) {
            int n = o->nod[k][i];
            o->un[n] += ui; o->vn[n] += vi; o->tn[n] += ti;
            o->rhon[n] += rhoi; o->rn[n] += ri;
        }
        for (int k = 0; k < 4; ++k) o->itest[o->nod[k][i]]++;
    }
    for (int gi = 1; gi <= o->nghosts; ++gi) {           /* viscous.f90:98-124 */
        int np = o->parent[gi], ng = o->ghost[gi], nt = o->type[gi], ns = o->side[gi];
        int n1 = o->nod[(ns - 1) & 3][np], n2 = o->nod[ns & 3][np];
        double rg = o->c[4][ng] / o->c[0][ng];
        o->un[n1] += o->u[ng]; o->vn[n1] += o->v[ng]; o->tn[n1] += o->t[ng];
        o->rhon[n1] += o->c[0][ng]; o->rn[n1] += rg;
        o->un[n2] += o->u[ng]; o->vn[n2] += o->v[ng]; o->tn[n2] += o->t[ng];
        o->rhon[n2] += o->c[0][ng]; o->rn[n2] += o->c[4][ng] / o->c[0][ng];
        if (nt == 4) {
            o->vn[n1] = 0.0; o->vn[n2] = 0.0;
        } else if (nt == 5) {
            o->un[n1] = 0.0; o->un[n2] = 0.0; o->vn[n1] = 0.0; o->vn[n2] = 0.0;
            o->rn[n1] = o->rtol * o->rtred; o->rn[n2] = o->rtol * o->rtred;
        }
        o->itest[n1]++; o->itest[n2]++;
    }
    for (int n = 1; n <= nn; ++n) {                      /* viscous.f90:51-59 */
        double cnt = (double)o->itest[n];
        o->un[n] = o->un[n] / cnt;
        o->vn[n] = o->vn[n] / cnt;
        o->tn[n] = o->tn[n] / cnt;
        o->rhon[n] = o->rhon[n] / cnt;
        o->rn[n] = o->rn[n] / cnt;
        o->rn[n] = dmax(o->rtol, o->rn[n]);
    }
    for (int i = 1; i <= o->neles; ++i) {                /* viscous.f90:62-70 */
        int n1 = o->nod[0][i], n2 = o->nod[1][i], n3 = o->nod[2][i], n4 = o->nod[3][i];
        viscous_face(o, i, 1, n1, n2, o->nc[0][i]);
        viscous_face(o, i, 2, n2, n3, o->nc[1][i]);
        viscous_face(o, i, 3, n3, n4, o->nc[2][i]);
        viscous_face(o, i, 4, n4, n1, o->nc[3][i]);
    }
}

/* source_terms.f90:24-104 */
void orc_source_terms(orc *o) {
    for (int i = 1; i <= o->nmax; ++i) { o->src3[i] = 0.0; o->src5[i] = 0.0; }
    const double sqcmu = sqrt(2.0 * sqrt(CMU_VAL));
    for (int i = 1; i <= o->neles; ++i) {
        int na = o->nc[0][i], nb = o->nc[1][i], nc_ = o->nc[2][i], nd = o->nc[3][i];
        double xa = o->xcc[na], xb = o->xcc[nb], xc = o->xcc[nc_], xd = o->xcc[nd];
        double ya = o->ycc[na], yb = o->ycc[nb], yc = o->ycc[nc_], yd = o->ycc[nd];
        double ua = o->u[na], ub = o->u[nb], uc = o->u[nc_], ud = o->u[nd];
        double va = o->v[na], vb = o->v[nb], vc = o->v[nc_], vd = o->v[nd];
        double ra = o->c[4][na] / o->c[0][na], rb = o->c[4][nb] / o->c[0][nb];
        double rc = o->c[4][nc_] / o->c[0][nc_], rd = o->c[4][nd] / o->c[0][nd];
        double qa = sqrt(ua * ua + va * va), qb = sqrt(ub * ub + vb * vb);
        double qc = sqrt(uc * uc + vc * vc), qd = sqrt(ud * ud + vd * vd);

        double dnr = (xa - xc) * (yb - yd) - (xb - xd) * (ya - yc);
        double dqx = ((qa - qc) * (yb - yd) - (qb - qd) * (ya - yc)) / dnr;
        double drx = ((ra - rc) * (yb - yd) - (rb - rd) * (ya - yc)) / dnr;
        double dux = ((ua - uc) * (yb - yd) - (ub - ud) * (ya - yc)) / dnr;
        double dvx = ((va - vc) * (yb - yd) - (vb - vd) * (ya - yc)) / dnr;
        double dqy = (-((qa - qc) * (xb - xd)) + (qb - qd) * (xa - xc)) / dnr;
        double dry = (-((ra - rc) * (xb - xd)) + (rb - rd) * (xa - xc)) / dnr;
        double duy = (-((ua - uc) * (xb - xd)) + (ub - ud) * (xa - xc)) / dnr;
        double dvy = (-((va - vc) * (xb - xd)) + (vb - vd) * (xa - xc)) / dnr;

        double rho = o->c[0][i];
        double ri = o->c[4][i] / rho;
        double amu = SUTH_COEF * sqrt(o->amol) * pow(o->t[i], 0.6);
        double chi = rho * ri / (CMU_VAL * amu);
        double fmu = tanh(ALPHA_TURB * chi * chi) / tanh(BETA_TURB * chi * chi);
        double fr2 = tanh(sqrt(chi)) / tanh(chi * sqcmu);
        double amut = fmu * rho * ri;
        double phi = dqx * drx + dqy * dry;
        double dee = 0.0;
        if (phi > 0.0) dee = drx * drx + dry * dry;
        double ratio = ipow(amut / dmax(amut, AMU_THRESH * amu), o->npseudo);
        if (o->iaxis == 1) {
            double vbyy = (o->ycc[i] < YAXIS_THRESH) ? dvy : o->v[i] / o->ycc[i];
            double divvel = dux + dvy + vbyy;
            double pk = amut / rho *
                        (dux * dux + dvy * dvy + vbyy * vbyy + dux * dux + dvy * dvy + vbyy * vbyy +
                         (duy + dvx) * (duy + dvx) - DIV23_VAL * divvel * divvel);
            pk = dmax(pk, 0.0);
            o->src5[i] = (rho * (CONST1_VAL - CONST2_VAL * fr2) * sqrt(ri * pk) -
                          rho * CONST3_VAL * dee) * ratio;
            double ttt = (amu + amut) * (vbyy + vbyy - DIV23_VAL * divvel);
            o->src3[i] = (o->p[i] - ttt) / o->ycc[i];
        } else {
            double divvel = dux + dvy;
            double pk = amut / rho *
                        (dux * dux + dvy * dvy + dux * dux + dvy * dvy + (duy + dvx) * (duy + dvx) -
                         DIV23_VAL * divvel * divvel);
            pk = dmax(pk, 0.0);
            o->src5[i] = (rho * (CONST1_VAL - CONST2_VAL * fr2) * sqrt(ri * pk) -
                          rho * CONST3_VAL * dee) * ratio;
        }
    }
}

/* residual.f90:23-62 */
void orc_residual(orc *o) {
    for (int i = 1; i <= o->neles; ++i) {
        int nb[4] = {o->nc[0][i], o->nc[1][i], o->nc[2][i], o->nc[3][i]};
        for (int m = 0; m < 5; ++m) {
            const double *F = o->f[m], *G = o->g[m];
            double fi = F[i], gi = G[i];
            double s = (fi + F[nb[0]]) * o->sc[0][i] + (gi + G[nb[0]]) * o->sc[1][i];
            s = s + (fi + F[nb[1]]) * o->sc[2][i];
            s = s + (gi + G[nb[1]]) * o->sc[3][i];
            s = s + (fi + F[nb[2]]) * o->sc[4][i];
            s = s + (gi + G[nb[2]]) * o->sc[5][i];
            s = s + (fi + F[nb[3]]) * o->sc[6][i];
            s = s + (gi + G[nb[3]]) * o->sc[7][i];
            o->r[m][i] = HALF * s;
        }
        o->r[4][i] = o->r[4][i] - o->vol[i] * o->src5[i];
        if (o->iaxis == 1) o->r[2][i] = o->r[2][i] - o->vol[i] * o->src3[i];
    }
}

/* residual.f90:67-83 */
void orc_update(orc *o, double alpha) {
    for (int i = 1; i <= o->neles; ++i) {
        double dtbyvol = o->dt[i] / o->vol[i];
        double a = alpha * dtbyvol;
        o->c[0][i] = o->cold[0][i] - a * (o->r[0][i] - o->d24[0][i]);
        o->c[1][i] = o->cold[1][i] - a * (o->r[1][i] - o->vis[0][i] - o->d24[1][i]);
        o->c[2][i] = o->cold[2][i] - a * (o->r[2][i] - o->vis[1][i] - o->d24[2][i]);
        o->c[3][i] = o->cold[3][i] - a * (o->r[3][i] - o->vis[2][i] - o->d24[3][i]);
        o->c[4][i] = o->cold[4][i] - a * CFLTURB_VAL * (o->r[4][i] - o->vis[3][i] - o->d24[4][i]);
        o->c[4][i] = o->c[0][i] * dmax(o->rtol, o->c[4][i] / o->c[0][i]);
    }
}

void orc_monitor(orc *o, double *l2, double *linf) {
    double s[5] = {0, 0, 0, 0, 0}, mx[5] = {0, 0, 0, 0, 0};
    for (int i = 1; i <= o->neles; ++i)
        for (int m = 0; m < 5; ++m) {
            double d = o->c[m][i] - o->cold[m][i];
            s[m] = s[m] + d * d;
            mx[m] = dmax(mx[m], fabs(d));
        }
    for (int m = 0; m < 5; ++m) {
        if (l2) l2[m] = sqrt(s[m] / (double)o->neles);
        if (linf) linf[m] = mx[m];
    }
    save_old(o);
}

/* turbnstwod.f90:95-168 */
void orc_iterate(orc *o, int32_t niter, double *l2, double *linf) {
    static const double alpha[4] = {0.25, 1.0 / 3.0, 0.5, 1.0};
    for (int it = 0; it < niter; ++it) {
        orc_timestep(o);
        orc_dissipation(o);
        for (int s = 0; s < 4; ++s) {
            orc_primitives(o);
            orc_apply_boundary(o);
            if (s == 0) orc_viscous(o);
            orc_source_terms(o);
            orc_residual(o);
            orc_update(o, alpha[s]);
        }
        orc_monitor(o, l2 ? l2 + 5 * (size_t)it : NULL, linf ? linf + 5 * (size_t)it : NULL);
    }
}

double *orc_array(orc *o, const char *nm, int32_t *len) {
    int nmax = o->nmax, ne = o->neles;
    struct { const char *n; double *a; int len; } tab[] = {
        {"U", o->u, nmax}, {"V", o->v, nmax}, {"T", o->t, nmax}, {"P", o->p, nmax},
        {"DT", o->dt, nmax}, {"ANU", o->anu, nmax}, {"ANURT", o->anurt, nmax},
        {"VISU", o->vis[0], ne}, {"VISV", o->vis[1], ne}, {"VIST", o->vis[2], ne},
        {"VISR", o->vis[3], ne}, {"SOURCE3", o->src3, ne}, {"SOURCE5", o->src5, ne},
        {"XCC", o->xcc, nmax}, {"YCC", o->ycc, nmax}, {"VOL", o->vol, nmax}, {"DL", o->dl, ne},
        {"UN", o->un, o->nods}, {"VN", o->vn, o->nods}, {"TN", o->tn, o->nods},
        {"RN", o->rn, o->nods}, {"RHON", o->rhon, o->nods},
        {"SC1X", o->sc[0], ne}, {"SC1Y", o->sc[1], ne}, {"SC2X", o->sc[2], ne},
        {"SC2Y", o->sc[3], ne}, {"SC3X", o->sc[4], ne}, {"SC3Y", o->sc[5], ne},
        {"SC4X", o->sc[6], ne}, {"SC4Y", o->sc[7], ne},
    };
    for (size_t k = 0; k < sizeof tab / sizeof tab[0]; ++k)
        if (!strcmp(nm, tab[k].n)) { *len = tab[k].len; return tab[k].a + 1; }
    static const char *fam[] = {"C", "COLD", "F", "G", "R", "D2Q", "D24"};
    for (int f = 0; f < 7; ++f) {
        size_t l = strlen(fam[f]);
        if (!strncmp(nm, fam[f], l) && strlen(nm) == l + 1 && nm[l] >= '1' && nm[l] <= '5') {
            int m = nm[l] - '1';
            double **arr[] = {o->c, o->cold, o->f, o->g, o->r, o->d2q, o->d24};
            *len = (f == 1 || f == 4 || f == 6) ? ne : nmax;
            return arr[f][m] + 1;
        }
    }
    *len = 0;
    return NULL;
}

double orc_scalar(orc *o, const char *nm) {
    if (!strcmp(nm, "RTOL")) return o->rtol;
    if (!strcmp(nm, "RINIT")) return o->rinit;
    if (!strcmp(nm, "DTMIN")) return o->dtmin;
    if (!strcmp(nm, "RR")) return o->rr;
    if (!strcmp(nm, "CV")) return o->cv;
    if (!strcmp(nm, "CP")) return o->cp;
    if (!strcmp(nm, "S0")) return o->s0;
    if (!strcmp(nm, "TINF")) return o->tinf;
    return NAN;
}
