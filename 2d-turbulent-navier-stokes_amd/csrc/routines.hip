// Routine-level kernels: one kernel (or a short chain) per reference routine,
// over reference-layout arrays.  They back the tdt_timestep .. tdt_monitor
// entry points, which a Fortran host calls in place of the reference's
// argument-less subroutines, and they are the unfused baseline the fused path
// (fused.hip) is checked against.
#include "ctx.hpp"
#include "reduce.hpp"

namespace tdt {

namespace {

__device__ __forceinline__ int gid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// geometry.f90:55-149 -----------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_geometry(Dev D, int iaxis) {
    int i = gid();
    if (i < D.n) {
        int4 nd = D.nod4[i];
        double x1 = D.xn[nd.x], x2 = D.xn[nd.y], x3 = D.xn[nd.z], x4 = D.xn[nd.w];
        double y1 = D.yn[nd.x], y2 = D.yn[nd.y], y3 = D.yn[nd.z], y4 = D.yn[nd.w];
        double xcc = (x1 + x2 + x3 + x4) * 0.25;
        double ycc = (y1 + y2 + y3 + y4) * 0.25;
        D.xcc[i] = xcc;
        D.ycc[i] = ycc;
        double dl12 = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
        double dl23 = sqrt((x2 - x3) * (x2 - x3) + (y2 - y3) * (y2 - y3));
        double dl34 = sqrt((x3 - x4) * (x3 - x4) + (y3 - y4) * (y3 - y4));
        double dl41 = sqrt((x4 - x1) * (x4 - x1) + (y4 - y1) * (y4 - y1));
        D.dl[i] = dmin(dmin(dmin(dl12, dl23), dl34), dl41);
        double *sc = D.sc;
        const int n = D.n;
        double vol;
        if (iaxis == 1) {
            sc[0 * n + i] = (y2 - y1) * (y1 + y2) * 0.5;
            sc[1 * n + i] = -((x2 - x1) * (y1 + y2) * 0.5);
            sc[2 * n + i] = (y3 - y2) * (y2 + y3) * 0.5;
            sc[3 * n + i] = -((x3 - x2) * (y2 + y3) * 0.5);
            sc[4 * n + i] = (y4 - y3) * (y3 + y4) * 0.5;
            sc[5 * n + i] = -((x4 - x3) * (y3 + y4) * 0.5);
            sc[6 * n + i] = (y1 - y4) * (y4 + y1) * 0.5;
            sc[7 * n + i] = -((x1 - x4) * (y4 + y1) * 0.5);
            vol = 0.5 * ((x3 - x1) * (y4 - y2) - (x4 - x2) * (y3 - y1)) * ycc;
        } else {
            sc[0 * n + i] = (y2 - y1);
            sc[1 * n + i] = -(x2 - x1);
            sc[2 * n + i] = (y3 - y2);
            sc[3 * n + i] = -(x3 - x2);
            sc[4 * n + i] = (y4 - y3);
            sc[5 * n + i] = -(x4 - x3);
            sc[6 * n + i] = (y1 - y4);
            sc[7 * n + i] = -(x1 - x4);
            vol = 0.5 * ((x3 - x1) * (y4 - y2) - (x4 - x2) * (y3 - y1));
        }
        if (vol <= 0.0) vol = -vol;
        D.vol[i] = vol;
    }
}

// ghost centroids mirrored through the face midpoint, VOL(NG) = VOL(NP)
__global__ void __launch_bounds__(kBlock) k_geometry_ghosts(Dev D) {
    int k = gid();
    if (k < D.g) {
        int par = D.gparent[k], sd = D.gside[k];
        int4 nd = D.nod4[par];
        int a = sd == 1 ? nd.x : sd == 2 ? nd.y : sd == 3 ? nd.z : nd.w;
        int b = sd == 1 ? nd.y : sd == 2 ? nd.z : sd == 3 ? nd.w : nd.x;
        double xf = 0.5 * (D.xn[a] + D.xn[b]);
        double yf = 0.5 * (D.yn[a] + D.yn[b]);
        D.xcc[D.n + k] = 2.0 * xf - D.xcc[par];
        D.ycc[D.n + k] = 2.0 * yf - D.ycc[par];
        D.vol[D.n + k] = D.vol[par];
    }
}

__global__ void k_set_scalar(double *p, double v) { *p = v; }

__global__ void __launch_bounds__(kBlock) k_cold_start(Dev D, double r0, double r1, double r2,
                                                       double r3, double r4) {
    int i = gid();
    if (i < D.n) {
        const double q[5] = {r0, r1, r2, r3, r4};
        for (int m = 0; m < 5; ++m) {
            D.c[m * D.nt + i] = q[m];
            D.cold[m * D.n + i] = q[m];
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_copy_cold(Dev D) {
    int i = gid();
    if (i < D.n)
        for (int m = 0; m < 5; ++m) D.cold[m * D.n + i] = D.c[m * D.nt + i];
}

// turbnstwod.f90:97-108, DTMIN as an order-independent atomic min -------------------
__global__ void __launch_bounds__(kBlock) k_timestep(Dev D, Consts<double> K) {
    int i = gid();
    double dtl = 100000.0;
    if (i < D.n) {
        const int nt = D.nt;
        double rho = D.c[i];
        double u = D.c[nt + i] / rho;
        double v = D.c[2 * nt + i] / rho;
        double vel = sqrt(u * u + v * v);
        double e = D.c[3 * nt + i] / rho;
        double t = (e - 0.5 * (u * u + v * v)) / K.cv;
        double p = rho * K.rr * t;
        double dt = K.cfl * D.dl[i] / (fabs(vel) + sqrt(K.gamma_rr * t));
        D.u[i] = u; D.v[i] = v; D.t[i] = t; D.p[i] = p; D.dt[i] = dt;
        dtl = dt;
    }
    block_atomic_min_pos(dtl, D.scal);
}

// turbnstwod.f90:110-114
__global__ void __launch_bounds__(kBlock) k_set_dt(Dev D) {
    int i = gid();
    if (i < D.n) D.dt[i] = D.scal[0];
}

// solver_state.f90:77-88
__global__ void __launch_bounds__(kBlock) k_ghost_copy1(Dev D) {
    int k = gid();
    if (k < D.g) {
        int par = D.gparent[k], gg = D.n + k;
        D.p[gg] = D.p[par];
        D.dt[gg] = D.dt[par];
        for (int m = 0; m < 5; ++m) D.c[m * D.nt + gg] = D.c[m * D.nt + par];
    }
}

// solver_state.f90:90-111
__global__ void __launch_bounds__(kBlock) k_dissip1(Dev D, double rtol) {
    int i = gid();
    if (i < D.n) {
        const int nt = D.nt;
        int4 nb = D.nc4[i];
        const double *P = D.p, *C5 = D.c + 4 * nt;
        double pie = D.p[i], c1i = D.c[i], c5i = C5[i];
        double anun = fabs(P[nb.x] - pie) + fabs(P[nb.y] - pie) + fabs(P[nb.z] - pie) +
                      fabs(P[nb.w] - pie);
        double anud = fabs(P[nb.x] + pie) + fabs(P[nb.y] + pie) + fabs(P[nb.z] + pie) +
                      fabs(P[nb.w] + pie);
        D.anu[i] = anun / anud;
        anun = fabs(C5[nb.x] - c5i) + fabs(C5[nb.y] - c5i) + fabs(C5[nb.z] - c5i) +
               fabs(C5[nb.w] - c5i);
        anud = fabs(C5[nb.x] + c5i) + fabs(C5[nb.y] + c5i) + fabs(C5[nb.z] + c5i) +
               fabs(C5[nb.w] + c5i);
        D.anurt[i] = anun / dmax(anud, 8.0 * c1i * rtol);
        for (int m = 0; m < 5; ++m) {
            const double *C = D.c + m * nt;
            D.d2q[m * nt + i] = C[nb.x] + C[nb.y] + C[nb.z] + C[nb.w] - 4.0 * C[i];
        }
    }
}

// solver_state.f90:113-118 (ANURT is not copied: ghost ANURT stays 0)
__global__ void __launch_bounds__(kBlock) k_ghost_copy2(Dev D) {
    int k = gid();
    if (k < D.g) {
        int par = D.gparent[k], gg = D.n + k;
        D.anu[gg] = D.anu[par];
        for (int m = 0; m < 5; ++m) D.d2q[m * D.nt + gg] = D.d2q[m * D.nt + par];
    }
}

// solver_state.f90:120-223
__global__ void __launch_bounds__(kBlock) k_dissip2(Dev D) {
    int i = gid();
    if (i < D.n) {
        const int nt = D.nt;
        int4 nb4 = D.nc4[i];
        const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
        double volbydt = D.vol[i] / D.dt[i];
        for (int m = 0; m < 5; ++m) {
            const double *sens = m < 4 ? D.anu : D.anurt;
            const double *C = D.c + m * nt, *Q = D.d2q + m * nt;
            double si = sens[i], d2 = 0.0, d4 = 0.0;
            for (int k = 0; k < 4; ++k) {
                int j = nb[k];
                double d2n, d4n;
                jst_coeffs(si, sens[j], volbydt, D.vol[j], D.dt[j], d2n, d4n);
                if (k == 0) {
                    d2 = d2n * (C[j] - C[i]);
                    d4 = d4n * (Q[j] - Q[i]);
                } else {
                    d2 = d2 + d2n * (C[j] - C[i]);
                    d4 = d4 + d4n * (Q[j] - Q[i]);
                }
            }
            D.d24[m * D.n + i] = d2 - d4;
        }
    }
}

// turbnstwod.f90:204-227
__global__ void __launch_bounds__(kBlock) k_prim(Dev D, Consts<double> K) {
    int i = gid();
    if (i < D.n) {
        const int nt = D.nt;
        Prim<double> s;
        prim_from_q(s, D.c[i], D.c[nt + i], D.c[2 * nt + i], D.c[3 * nt + i], D.c[4 * nt + i], K);
        D.u[i] = s.u; D.v[i] = s.v; D.t[i] = s.t; D.p[i] = s.p;
        for (int m = 0; m < 5; ++m) {
            D.f[m * nt + i] = s.f[m];
            D.gf[m * nt + i] = s.g[m];
        }
    }
}

__device__ __forceinline__ void load_prim(const Dev &D, int id, Prim<double> &s) {
    const int nt = D.nt;
    for (int m = 0; m < 5; ++m) {
        s.c[m] = D.c[m * nt + id];
        s.f[m] = D.f[m * nt + id];
        s.g[m] = D.gf[m * nt + id];
    }
    s.u = D.u[id]; s.v = D.v[id]; s.t = D.t[id]; s.p = D.p[id];
}

// boundary.f90:24-48: one thread per ghost, reading only interior cells
__global__ void __launch_bounds__(kBlock) k_bc(Dev D, Consts<double> K) {
    int k = gid();
    if (k < D.g) {
        int par = D.gparent[k], opp = D.gopp[k], gg = D.n + k;
        Prim<double> pp, s;
        load_prim(D, par, pp);
        const int nt = D.nt;
        if (D.gtype[k] == 6) {
            const int f = D.gside[k] - 1, n = D.n;
            bc6_state(s, pp, D.sc[(2 * f) * n + par], D.sc[(2 * f + 1) * n + par],
                      fabs(D.ycc[gg] + D.ycc[par]), D.g6ptr[k], D.g6ptr[k + 1], D.g6cell,
                      D.g6dely,
                      [&](int c, double &c1, double &c2, double &yc) {
                          c1 = D.c[c]; c2 = D.c[nt + c]; yc = D.ycc[c];
                      },
                      K);
        } else {
            ghost_state(s, D.gtype[k], pp, D.u[opp], D.v[opp], D.t[opp], D.p[opp], D.ycc[gg], K);
        }
        for (int m = 0; m < 5; ++m) {
            D.c[m * nt + gg] = s.c[m];
            D.f[m * nt + gg] = s.f[m];
            D.gf[m * nt + gg] = s.g[m];
        }
        D.u[gg] = s.u; D.v[gg] = s.v; D.t[gg] = s.t; D.p[gg] = s.p;
    }
}

// viscous.f90:40-59, 77-124 as a node-centric gather (no atomics)
__global__ void __launch_bounds__(kBlock) k_nodeavg(Dev D, Consts<double> K) {
    int v = gid();
    if (v < D.m) {
        const int nt = D.nt;
        int4 ov = D.novr[v];
        int j0 = D.nptr[v], j1 = D.nptr[v + 1];
        double au = 0.0, av = 0.0, at = 0.0, arho = 0.0, ar = ov.w ? K.rtol_rtred : 0.0;
        for (int j = j0; j < j1; ++j) {
            int id = D.nlist[j];
            double rho = D.c[id], rr = D.c[4 * nt + id] / rho;
            if (j >= ov.x) au += D.u[id];
            if (j >= ov.y) av += D.v[id];
            at += D.t[id];
            arho += rho;
            if (j >= ov.z) ar += rr;
        }
        double cnt = (double)(j1 - j0);
        const int m = D.m;
        D.nodeavg[v] = au / cnt;
        D.nodeavg[m + v] = av / cnt;
        D.nodeavg[2 * m + v] = at / cnt;
        D.nodeavg[3 * m + v] = arho / cnt;
        D.nodeavg[4 * m + v] = dmax(K.rtol, ar / cnt);
    }
}

// POSTPROCESS (solver_state.f90:230-283): node averages of P, U, V, T over the
// cells around each node and the ghosts on its faces, in the reference's order
// (the node lists: cells by id, then ghosts in bc.in order; no overrides), then
// the tec_plot.dat columns U, V, P [bar], RHO, MACH, T.  out: [6][m].
__global__ void __launch_bounds__(kBlock) k_postprocess(Dev D, Consts<double> K, double *out) {
    int v = gid();
    if (v < D.m) {
        const int j0 = D.nptr[v], j1 = D.nptr[v + 1];
        double pn = 0.0, un = 0.0, vn = 0.0, tn = 0.0;
        for (int j = j0; j < j1; ++j) {
            const int id = D.nlist[j];
            pn = pn + D.p[id];
            un = un + D.u[id];
            vn = vn + D.v[id];
            tn = tn + D.t[id];
        }
        const double df = (double)(j1 - j0);
        pn = pn / df; un = un / df; vn = vn / df; tn = tn / df;
        const int m = D.m;
        out[v] = un;
        out[m + v] = vn;
        out[2 * m + v] = pn * kPaToBar;
        out[3 * m + v] = pn / (K.rr * tn);
        out[4 * m + v] = sqrt(un * un + vn * vn) / sqrt(K.gamma_rr * tn);
        out[5 * m + v] = tn;
    }
}

// viscous.f90:62-70, 130-217
__global__ void __launch_bounds__(kBlock) k_viscous(Dev D, Consts<double> K) {
    int i = gid();
    if (i < D.n) {
        const int nt = D.nt, n = D.n, m = D.m;
        int4 nd4 = D.nod4[i], nb4 = D.nc4[i];
        const int nd[4] = {nd4.x, nd4.y, nd4.z, nd4.w};
        const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
        double xa = D.xcc[i], ya = D.ycc[i];
        double ua = D.u[i], va = D.v[i], ta = D.t[i], rhoa = D.c[i], ra = D.c[4 * nt + i] / rhoa;
        const double *N = D.nodeavg;
        double vu = 0.0, vv = 0.0, vt = 0.0, vr = 0.0;
        for (int f = 0; f < 4; ++f) {
            int b = nd[f], d = nd[(f + 1) & 3], c = nb[f];
            double rhoc = D.c[c];
            ViscFace<double> o = viscous_face<double>(
                D.sc[(2 * f) * n + i], D.sc[(2 * f + 1) * n + i], xa, ya, D.xn[b], D.yn[b],
                D.xcc[c], D.ycc[c], D.xn[d], D.yn[d], ua, va, ta, rhoa, ra, N[b], N[m + b],
                N[2 * m + b], N[3 * m + b], N[4 * m + b], D.u[c], D.v[c], D.t[c], rhoc,
                D.c[4 * nt + c] / rhoc, N[d], N[m + d], N[2 * m + d], N[3 * m + d], N[4 * m + d], K);
            vu = vu + o.u0 + o.u1;
            vv = vv + o.v0 + o.v1;
            vt = vt + o.t0 + o.t1;
            vr = vr + o.r0 + o.r1;
        }
        D.vis[i] = vu; D.vis[n + i] = vv; D.vis[2 * n + i] = vt; D.vis[3 * n + i] = vr;
    }
}

// source_terms.f90:24-104
__global__ void __launch_bounds__(kBlock) k_source(Dev D, Consts<double> K) {
    int i = gid();
    if (i < D.n) {
        const int nt = D.nt;
        int4 nb = D.nc4[i];
        const double *C1 = D.c, *C5 = D.c + 4 * nt;
        double s3;
        auto nbr = [&](int j) {
            return src_nbr<double>(D.xcc[j], D.ycc[j], D.u[j], D.v[j], C1[j], C5[j]);
        };
        double s5 = source_terms<double>(nbr(nb.x), nbr(nb.y), nbr(nb.z), nbr(nb.w), C1[i], C5[i],
                                         D.t[i], D.p[i], D.v[i], D.ycc[i], s3, K);
        D.src5[i] = s5;
        D.src3[i] = s3;
    }
}

// residual.f90:23-62
__global__ void __launch_bounds__(kBlock) k_residual(Dev D, int iaxis) {
    int i = gid();
    if (i < D.n) {
        const int nt = D.nt, n = D.n;
        int4 nb4 = D.nc4[i];
        const int nb[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
        double sc[8];
        for (int k = 0; k < 8; ++k) sc[k] = D.sc[k * n + i];
        for (int m = 0; m < 5; ++m) {
            const double *F = D.f + m * nt, *G = D.gf + m * nt;
            double fi = F[i], gi = G[i];
            double s = (fi + F[nb[0]]) * sc[0] + (gi + G[nb[0]]) * sc[1];
            for (int k = 1; k < 4; ++k) {
                s = s + (fi + F[nb[k]]) * sc[2 * k];
                s = s + (gi + G[nb[k]]) * sc[2 * k + 1];
            }
            D.r[m * n + i] = 0.5 * s;
        }
        D.r[4 * n + i] = D.r[4 * n + i] - D.vol[i] * D.src5[i];
        if (iaxis == 1) D.r[2 * n + i] = D.r[2 * n + i] - D.vol[i] * D.src3[i];
    }
}

// residual.f90:67-83
__global__ void __launch_bounds__(kBlock) k_update(Dev D, double alpha, double rtol) {
    int i = gid();
    if (i < D.n) {
        const int nt = D.nt, n = D.n;
        double a = alpha * (D.dt[i] / D.vol[i]);
        double *C = D.c;
        const double *CO = D.cold, *R = D.r, *V = D.vis, *Q = D.d24;
        double c1 = CO[i] - a * (R[i] - Q[i]);
        double c2 = CO[n + i] - a * (R[n + i] - V[i] - Q[n + i]);
        double c3 = CO[2 * n + i] - a * (R[2 * n + i] - V[n + i] - Q[2 * n + i]);
        double c4 = CO[3 * n + i] - a * (R[3 * n + i] - V[2 * n + i] - Q[3 * n + i]);
        double c5 = CO[4 * n + i] - a * kCflTurb * (R[4 * n + i] - V[3 * n + i] - Q[4 * n + i]);
        c5 = c1 * dmax(rtol, c5 / c1);
        C[i] = c1; C[nt + i] = c2; C[2 * nt + i] = c3; C[3 * nt + i] = c4; C[4 * nt + i] = c5;
    }
}

// L2 / Linf partials of C - COLD per block, optionally COLD <- C
__global__ void __launch_bounds__(kBlock) k_monitor_partial(Dev D, int save_old) {
    int i = gid();
    double sq[5] = {0, 0, 0, 0, 0}, mx[5] = {0, 0, 0, 0, 0};
    int bad = 0;
    if (i < D.n) {
        for (int m = 0; m < 5; ++m) {
            double cn = D.c[m * D.nt + i];
            double d = cn - D.cold[m * D.n + i];
            sq[m] = d * d;
            mx[m] = fabs(d);
            bad |= !isfinite(cn);
            if (save_old) D.cold[m * D.n + i] = cn;
        }
    }
    block_partials_write(sq, mx, bad, D.part + (size_t)blockIdx.x * kPartStride);
}

__global__ void __launch_bounds__(kBlock) k_monitor_final(const double *part, int nblk, int n,
                                                          double *out, int *flag) {
    reduce_partials(part, nblk, n, out, flag);
}

}  // namespace

#define LAUNCH(kern, count, ...)                                                   \
    do {                                                                           \
        if ((count) > 0) {                                                         \
            hipLaunchKernelGGL(kern, dim3(nblocks(count)), dim3(kBlock), 0, c->stream, \
                               __VA_ARGS__);                                       \
            hipError_t e_ = hipGetLastError();                                     \
            if (e_ != hipSuccess) return e_;                                       \
        }                                                                          \
    } while (0)

hipError_t launch_geometry(tdt_ctx *c, int iaxis) {
    LAUNCH(k_geometry, c->D.n, c->D, iaxis);
    LAUNCH(k_geometry_ghosts, c->D.g, c->D);
    c->sc_from_nodes = true;
    return hipSuccess;
}

hipError_t launch_timestep(tdt_ctx *c) {
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, c->stream, c->D.scal, 100000.0);
    LAUNCH(k_timestep, c->D.n, c->D, c->K);
    if (c->K.local == 0) LAUNCH(k_set_dt, c->D.n, c->D);
    return hipGetLastError();
}

hipError_t launch_dissipation(tdt_ctx *c) {
    LAUNCH(k_ghost_copy1, c->D.g, c->D);
    LAUNCH(k_dissip1, c->D.n, c->D, c->K.rtol);
    LAUNCH(k_ghost_copy2, c->D.g, c->D);
    LAUNCH(k_dissip2, c->D.n, c->D);
    return hipSuccess;
}

hipError_t launch_primitives(tdt_ctx *c) {
    LAUNCH(k_prim, c->D.n, c->D, c->K);
    return hipSuccess;
}

hipError_t launch_boundary(tdt_ctx *c) {
    LAUNCH(k_bc, c->D.g, c->D, c->K);
    return hipSuccess;
}

hipError_t launch_viscous(tdt_ctx *c) {
    LAUNCH(k_nodeavg, c->D.m, c->D, c->K);
    LAUNCH(k_viscous, c->D.n, c->D, c->K);
    return hipSuccess;
}

hipError_t launch_source(tdt_ctx *c) {
    LAUNCH(k_source, c->D.n, c->D, c->K);
    return hipSuccess;
}

hipError_t launch_residual(tdt_ctx *c) {
    LAUNCH(k_residual, c->D.n, c->D, c->K.iaxis);
    return hipSuccess;
}

hipError_t launch_update(tdt_ctx *c, double alpha) {
    LAUNCH(k_update, c->D.n, c->D, alpha, c->K.rtol);
    return hipSuccess;
}

hipError_t launch_monitor(tdt_ctx *c, int save_old, double *out) {
    int nb = (int)nblocks(c->D.n);
    LAUNCH(k_monitor_partial, c->D.n, c->D, save_old);
    hipLaunchKernelGGL(k_monitor_final, dim3(1), dim3(kBlock), 0, c->stream, c->D.part, nb,
                       c->D.n, out, c->D.flags);
    return hipGetLastError();
}

hipError_t launch_cold_start(tdt_ctx *c, double r0, double r1, double r2, double r3, double r4) {
    LAUNCH(k_cold_start, c->D.n, c->D, r0, r1, r2, r3, r4);
    return hipSuccess;
}

hipError_t launch_postprocess(tdt_ctx *c, double *out) {
    LAUNCH(k_postprocess, c->D.m, c->D, c->K, out);
    return hipSuccess;
}

hipError_t launch_copy_cold(tdt_ctx *c) {
    LAUNCH(k_copy_cold, c->D.n, c->D);
    return hipSuccess;
}

}  // namespace tdt
