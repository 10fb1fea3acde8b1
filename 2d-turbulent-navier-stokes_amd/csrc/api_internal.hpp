// Helpers shared by the translation units of the C ABI (api.cpp: contexts,
// state, routine entry points, arrays; api_batch.cpp: batches of cases;
// api_adjoint.cpp: the residual, its adjoint and parameter/node updates;
// api_adjoint_iter.cpp: the adjoint of whole iterations).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "ctx.hpp"
#include "tdt.h"

namespace tdt {
namespace api {

// message of the last failed context creation (tdt_last_error(NULL))
extern thread_local std::string g_create_error;

#define CHECK_CTX(c)                   \
    do {                               \
        if (!(c)) return TDT_ERR_ARG;  \
    } while (0)

// Entry points of one case: a batch context runs its cases through the
// tdt_batch_* calls only.
#define CHECK_SINGLE(c)                                                         \
    do {                                                                        \
        if (!(c)) return TDT_ERR_ARG;                                           \
        if ((c)->batch) {                                                       \
            (c)->err = std::string(__func__) + ": not available on a batch context"; \
            return TDT_ERR_UNSUPPORTED;                                         \
        }                                                                       \
    } while (0)

inline int hip_fail(tdt_ctx *c, hipError_t e, const char *what) {
    c->err = std::string(what) + ": " + hipGetErrorString(e);
    return TDT_ERR_DEVICE;
}

#define HIPCALL(c, call, what)                                  \
    do {                                                        \
        hipError_t e__ = (call);                                \
        if (e__ != hipSuccess) return hip_fail((c), e__, what); \
    } while (0)

template <class T>
T *dev_alloc(tdt_ctx *c, size_t count, bool zero = true) {
    void *p = nullptr;
    size_t bytes = count * sizeof(T) + 256;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    if (zero && hipMemset(p, 0, bytes) != hipSuccess) {
        hipFree(p);
        return nullptr;
    }
    c->allocs.push_back(p);
    return static_cast<T *>(p);
}

template <class R>
Consts<R> make_consts(const tdt_params &P) {
    // thermodynamics as in turbnstwod.f90:37-42; derived groupings keep the
    // reference's evaluation order so they are bit-identical to it
    double rr = 8314.0 / P.amol;
    double cv = rr / (P.gama - 1.0);
    double cp = cv * P.gama;
    double rho0 = P.p0 / (rr * P.t0);
    double s0 = P.p0 / std::pow(rho0, P.gama);
    double tinf = P.t0 / (1.0 + 0.5 * (P.gama - 1.0) * P.aminf * P.aminf);
    Consts<R> K{};
    K.gama = (R)P.gama; K.cfl = (R)P.cfl; K.prandtl = (R)P.prandtl; K.prandtlt = (R)P.prandtlt;
    K.t0 = (R)P.t0; K.pamb = (R)P.pamb; K.pinf = (R)P.pinf; K.deltabl = (R)P.deltabl;
    K.rr = (R)rr; K.cv = (R)cv; K.cp = (R)cp; K.s0 = (R)s0; K.tinf = (R)tinf;
    K.suth_sqrt_amol = (R)(tdt::kSuth * std::sqrt(P.amol));
    K.inv_1mg = (R)(1.0 / (1.0 - P.gama));
    K.sqcmu = (R)std::sqrt(2.0 * std::sqrt(tdt::kCmu));
    K.bl_u = (R)(2.0 * P.aminf * std::sqrt(P.gama * rr * tinf));
    K.gamma_rr = (R)(P.gama * rr);
    K.tflame = (R)P.tflame; K.ancomb = (R)P.ancomb; K.acomb = (R)P.acomb;
    K.rhoprop = (R)P.rhoprop;
    K.gth_08 = (R)std::pow(tdt::kGth, 0.8);
    K.local = P.local; K.npseudo = P.npseudo; K.iaxis = P.iaxis;
    K.pamb_set = P.pamb >= 0.0;
    return K;
}

// RTOL = RINIT * RTRED and RTOL * RTRED in both precisions' constants.
void set_rinit(tdt_ctx *c, double rinit);
int sync(tdt_ctx *c);
// flow.in values out of range: false, with g_create_error set
bool params_ok(const tdt_params &P);
// Device side of a context whose mesh is prepared (c->pm): stream, uploads,
// work arrays and the geometry (geometry.f90:55-149).  On failure the caller
// destroys c.
int init_device(tdt_ctx *c, const tdt_params *P);
// Uniform initial state of a cold start (turbnstwod.f90:53-71) with the host's
// libm, as the reference: q = C1..C5, *rinit = RINIT.
void cold_state(const tdt_params &P, const Consts<double> &K, double q[5], double *rinit);

// Residual adjoint buffers (api_adjoint.cpp), allocated on first use.
int adj_setup(tdt_ctx *c);
// [5][neles] in grid.in cell order (host) <-> [5][n] internal order (device)
int upload_cells(tdt_ctx *c, const double *src, double *dst);
int download_cells(tdt_ctx *c, const double *src, double *dst);

}  // namespace api
}  // namespace tdt
