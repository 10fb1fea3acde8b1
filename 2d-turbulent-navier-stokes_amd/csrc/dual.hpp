// Forward-mode dual numbers for the residual's local Jacobian blocks.
//
// Dual<N> carries a value and N directional derivatives.  Its value part is
// computed with exactly the FP64 operations the double code performs (the same
// IEEE +, -, *, /, sqrt and the tm:: functions), so a residual evaluated in
// Dual<N> has the double path's value bit for bit; the derivative parts follow
// the chain rule of each operation.  Comparisons look at the value (a branch
// taken in the value computation is the one differentiated).
#pragma once

#include <hip/hip_runtime.h>

#include "phys.hpp"
#include "tmath.hpp"

namespace tdt {

template <int N>
struct Dual {
    double v;
    double d[N];
    Dual() = default;
    __host__ __device__ Dual(double x) : v(x) {
#pragma unroll
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
};

#define TDT_DUAL_CMP(OP)                                                                          \
    template <int N>                                                                              \
    __device__ __forceinline__ bool operator OP(const Dual<N> &a, const Dual<N> &b) {             \
        return a.v OP b.v;                                                                        \
    }                                                                                             \
    template <int N>                                                                              \
    __device__ __forceinline__ bool operator OP(const Dual<N> &a, double b) { return a.v OP b; }  \
    template <int N>                                                                              \
    __device__ __forceinline__ bool operator OP(double a, const Dual<N> &b) { return a OP b.v; }
TDT_DUAL_CMP(<)
TDT_DUAL_CMP(>)
TDT_DUAL_CMP(<=)
TDT_DUAL_CMP(>=)
TDT_DUAL_CMP(==)
TDT_DUAL_CMP(!=)
#undef TDT_DUAL_CMP

template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a) {
    Dual<N> r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}

// (a/b)' = (a' - (a/b) b') / b
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r;
    r.v = a.v / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}

template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, double b) { return a + Dual<N>(b); }
template <int N>
__device__ __forceinline__ Dual<N> operator+(double a, const Dual<N> &b) { return Dual<N>(a) + b; }
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, double b) { return a - Dual<N>(b); }
template <int N>
__device__ __forceinline__ Dual<N> operator-(double a, const Dual<N> &b) { return Dual<N>(a) - b; }
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, double b) { return a * Dual<N>(b); }
template <int N>
__device__ __forceinline__ Dual<N> operator*(double a, const Dual<N> &b) { return Dual<N>(a) * b; }
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, double b) { return a / Dual<N>(b); }
template <int N>
__device__ __forceinline__ Dual<N> operator/(double a, const Dual<N> &b) { return Dual<N>(a) / b; }

template <int N>
__device__ __forceinline__ Dual<N> &operator*=(Dual<N> &a, const Dual<N> &b) { return a = a * b; }
template <int N>
__device__ __forceinline__ Dual<N> &operator+=(Dual<N> &a, const Dual<N> &b) { return a = a + b; }
template <int N>
__device__ __forceinline__ Dual<N> &operator-=(Dual<N> &a, const Dual<N> &b) { return a = a - b; }

// |a|: the derivative's sign follows the value (+ at 0, as the value's branch)
template <int N>
__device__ __forceinline__ Dual<N> fabs(const Dual<N> &a) { return a.v < 0.0 ? -a : a; }

// sqrt' = 1 / (2 sqrt); at exactly 0 the derivatives are set to 0: there the
// argument is a square-like quantity (the turbulence production of a uniform
// state) whose own derivative is 0, so sqrt of it behaves like |x|, which has no
// derivative at 0 and whose central differences are 0
template <int N>
__device__ __forceinline__ Dual<N> tsqrt(const Dual<N> &a) {
    Dual<N> r;
    r.v = sqrt(a.v);
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = r.v > 0.0 ? a.d[k] / (r.v + r.v) : 0.0;
    return r;
}

// tanh' = 1 - tanh^2
template <int N>
__device__ __forceinline__ Dual<N> ttanh(const Dual<N> &a) {
    Dual<N> r;
    r.v = tm::tanh(a.v);
    const double s = 1.0 - r.v * r.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = s * a.d[k];
    return r;
}

// (a^b)' = a^b (b' log a + b a' / a); b' is dropped when it is zero (the
// exponents here are constants), so a may be 0 with a constant exponent
template <int N>
__device__ __forceinline__ Dual<N> tpow(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r;
    r.v = tm::pow(a.v, b.v);
    bool vb = false;
#pragma unroll
    for (int k = 0; k < N; ++k) vb = vb || b.d[k] != 0.0;
    const double la = vb ? tm::log_pos(a.v) : 0.0;
    const double g = a.v != 0.0 ? r.v * b.v / a.v : 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = g * a.d[k] + (vb ? r.v * la * b.d[k] : 0.0);
    return r;
}

template <int N>
struct Recip<Dual<N>> {
    Dual<N> b;
    __device__ __forceinline__ explicit Recip(const Dual<N> &den) : b(den) {}
    __device__ __forceinline__ Dual<N> div(const Dual<N> &a) const { return a / b; }
};

// Dual<N> flow constants with zero derivatives.
template <int N>
__device__ __forceinline__ Consts<Dual<N>> lift(const Consts<double> &K) {
    Consts<Dual<N>> D;
    D.gama = K.gama; D.cfl = K.cfl; D.prandtl = K.prandtl; D.prandtlt = K.prandtlt;
    D.t0 = K.t0; D.pamb = K.pamb; D.pinf = K.pinf; D.deltabl = K.deltabl;
    D.rr = K.rr; D.cv = K.cv; D.cp = K.cp; D.s0 = K.s0; D.tinf = K.tinf;
    D.rtol = K.rtol; D.rtol_rtred = K.rtol_rtred;
    D.suth_sqrt_amol = K.suth_sqrt_amol; D.inv_1mg = K.inv_1mg; D.sqcmu = K.sqcmu;
    D.bl_u = K.bl_u; D.gamma_rr = K.gamma_rr;
    D.tflame = K.tflame; D.ancomb = K.ancomb; D.acomb = K.acomb; D.rhoprop = K.rhoprop;
    D.gth_08 = K.gth_08;
    D.local = K.local; D.npseudo = K.npseudo; D.iaxis = K.iaxis; D.pamb_set = K.pamb_set;
    return D;
}

}  // namespace tdt
