// Quad primitives: four consecutive lanes of a wavefront ("a quad") cooperate
// on one cell, lane f owning face f.  Values move inside a quad with DPP
// quad_perm moves (one VALU op per 32 bits, no LDS).  Every call must be made
// with all four lanes of the quad active.
#pragma once

#include <hip/hip_runtime.h>

namespace tdt {

// value held by lane K of this lane's quad
template <int K>
__device__ __forceinline__ int qget(int x) {
    return __builtin_amdgcn_mov_dpp(x, K * 0x55, 0xF, 0xF, false);
}
template <int K>
__device__ __forceinline__ float qget(float x) {
    return __int_as_float(qget<K>(__float_as_int(x)));
}
template <int K>
__device__ __forceinline__ double qget(double x) {
    int lo = qget<K>(__double2loint(x));
    int hi = qget<K>(__double2hiint(x));
    return __hiloint2double(hi, lo);
}

// value held by the lane across the cell (face f <-> face f^2): quad_perm(2,3,0,1)
__device__ __forceinline__ int qopp(int x) { return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false); }
__device__ __forceinline__ float qopp(float x) { return __int_as_float(qopp(__float_as_int(x))); }
__device__ __forceinline__ double qopp(double x) {
    int lo = qopp(__double2loint(x));
    int hi = qopp(__double2hiint(x));
    return __hiloint2double(hi, lo);
}

// ((x0 + x1) + x2) + x3 over the quad, lane order: the reference's face order
template <class R>
__device__ __forceinline__ R qsum(R x) {
    R s = qget<0>(x);
    s = s + qget<1>(x);
    s = s + qget<2>(x);
    s = s + qget<3>(x);
    return s;
}

// ((((((a0 + b0) + a1) + b1) + a2) + b2) + a3) + b3: two terms per face, in order
template <class R>
__device__ __forceinline__ R qsum2(R a, R b) {
    R s = qget<0>(a) + qget<0>(b);
    s = s + qget<1>(a);
    s = s + qget<1>(b);
    s = s + qget<2>(a);
    s = s + qget<2>(b);
    s = s + qget<3>(a);
    s = s + qget<3>(b);
    return s;
}

// |x0 - c| + |x1 - c| + ... and |x0 + c| + ..., lane order
template <class R>
__device__ __forceinline__ void qabs_sums(R x, R c, R &dsum, R &ssum) {
    R x0 = qget<0>(x), x1 = qget<1>(x), x2 = qget<2>(x), x3 = qget<3>(x);
    dsum = fabs(x0 - c) + fabs(x1 - c) + fabs(x2 - c) + fabs(x3 - c);
    ssum = fabs(x0 + c) + fabs(x1 + c) + fabs(x2 + c) + fabs(x3 + c);
}

// pick lane-indexed value: v[lane] with lane in 0..3, without dynamic register indexing
template <class R>
__device__ __forceinline__ R sel4(int lane, R v0, R v1, R v2, R v3) {
    return lane == 0 ? v0 : lane == 1 ? v1 : lane == 2 ? v2 : v3;
}

}  // namespace tdt
