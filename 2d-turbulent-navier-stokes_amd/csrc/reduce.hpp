// Deterministic reductions: 64-lane wavefront DPP butterflies, then a
// fixed-order combination of the block's waves.  Sums are reproducible run to
// run because the grid shape fixes the order; min / max are order-independent.
#pragma once

#include <hip/hip_runtime.h>

namespace tdt {

constexpr int kWave = 64;
constexpr int kPartStride = 11;  // per block: 5 sums of squares, 5 maxima, 1 flag

// Value of x in lane (i - K) mod 16 of this lane's row of 16 (DPP row_ror:K,
// a VALU move: no LDS round trip).
template <int K>
__device__ __forceinline__ double row_ror(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x120 + K, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x120 + K, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_value(double x, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), lane),
                            __builtin_amdgcn_readlane(__double2loint(x), lane));
}

// Wavefront reductions (all 64 lanes active): a rotate-and-combine butterfly
// inside each row of 16 lanes, then the four row results (lanes 0, 16, 32, 48)
// in row order.  The result is uniform; the order is fixed, so sums repeat
// bit for bit.
template <class Op>
__device__ __forceinline__ double wave_reduce(double x, Op op) {
    x = op(x, row_ror<8>(x));
    x = op(x, row_ror<4>(x));
    x = op(x, row_ror<2>(x));
    x = op(x, row_ror<1>(x));
    return op(op(op(lane_value(x, 0), lane_value(x, 16)), lane_value(x, 32)), lane_value(x, 48));
}

__device__ __forceinline__ double wave_sum(double x) {
    return wave_reduce(x, [](double a, double b) { return a + b; });
}

__device__ __forceinline__ double wave_max(double x) {
    return wave_reduce(x, [](double a, double b) { return b > a ? b : a; });
}

__device__ __forceinline__ double wave_min(double x) {
    return wave_reduce(x, [](double a, double b) { return b < a ? b : a; });
}

// Block minimum of non-negative values folded into *slot with an atomic min on
// the IEEE bit pattern (monotonic for non-negative doubles).
__device__ __forceinline__ void block_atomic_min_pos(double x, double *slot) {
    __shared__ double sh[16];
    x = wave_min(x);
    int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) sh[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sh[0];
        int nw = (blockDim.x + kWave - 1) / kWave;
        for (int q = 1; q < nw; ++q) m = sh[q] < m ? sh[q] : m;
        if (m >= 0.0)
            atomicMin(reinterpret_cast<unsigned long long *>(slot),
                      (unsigned long long)__double_as_longlong(m));
    }
}

// Combination of partial record field q: a sum for the sums of squares, a
// maximum (keeping a, as wave_max does, unless b is larger) for the rest.
__device__ __forceinline__ double part_op(int q, double a, double b) {
    return q < 5 ? a + b : (b > a ? b : a);
}

// Writes this block's partial sums of squares, maxima and non-finite flag.
// The eleven row butterflies run side by side (no value waits on another's
// DPP chain); each row's result goes through LDS, and one thread per field
// combines rows 0..3 of wave 0, 1, ... in that order: the same operations, in
// the same order, as wave_sum / wave_max followed by a fold over the waves.
__device__ __forceinline__ void block_partials_write(const double *sq, const double *mx,
                                                     int bad, double *out) {
    __shared__ double sh[16 * 4][kPartStride];   // [wave * 4 + row of 16 lanes][field]
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    const int nw = (blockDim.x + kWave - 1) / kWave;
    double x[kPartStride];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        x[m] = sq[m];
        x[5 + m] = mx[m];
    }
    x[10] = (double)bad;
#pragma unroll
    for (int q = 0; q < kPartStride; ++q) x[q] = part_op(q, x[q], row_ror<8>(x[q]));
#pragma unroll
    for (int q = 0; q < kPartStride; ++q) x[q] = part_op(q, x[q], row_ror<4>(x[q]));
#pragma unroll
    for (int q = 0; q < kPartStride; ++q) x[q] = part_op(q, x[q], row_ror<2>(x[q]));
#pragma unroll
    for (int q = 0; q < kPartStride; ++q) x[q] = part_op(q, x[q], row_ror<1>(x[q]));
    if ((lane & 15) == 0) {
#pragma unroll
        for (int q = 0; q < kPartStride; ++q) sh[w * 4 + (lane >> 4)][q] = x[q];
    }
    __syncthreads();
    if (threadIdx.x < kPartStride) {
        const int q = threadIdx.x;
        double acc = 0.0;
        for (int k = 0; k < nw; ++k) {
            double r = sh[4 * k][q];
            r = part_op(q, r, sh[4 * k + 1][q]);
            r = part_op(q, r, sh[4 * k + 2][q]);
            r = part_op(q, r, sh[4 * k + 3][q]);
            acc = k == 0 ? r : part_op(q, acc, r);
        }
        out[q] = acc;
    }
}

// One block folds nblk partial records in a fixed order:
// out[0..4] = sqrt(sum / n), out[5..9] = max, *flag |= any non-finite.
__device__ __forceinline__ void reduce_partials(const double *part, int nblk, int n,
                                                double *out, int *flag) {
    __shared__ double sh[16 * 4][kPartStride];   // [wave * 4 + row of 16 lanes][field]
    double acc[kPartStride];
    for (int q = 0; q < kPartStride; ++q) acc[q] = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        const double *p = part + (size_t)b * kPartStride;
        for (int q = 0; q < 5; ++q) acc[q] = acc[q] + p[q];
        for (int q = 5; q < kPartStride; ++q) acc[q] = p[q] > acc[q] ? p[q] : acc[q];
    }
    // as in block_partials_write: row butterflies side by side, rows 0..3 of
    // each wave combined in order through LDS
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    const int nw = (blockDim.x + kWave - 1) / kWave;
#pragma unroll
    for (int q = 0; q < kPartStride; ++q) acc[q] = part_op(q, acc[q], row_ror<8>(acc[q]));
#pragma unroll
    for (int q = 0; q < kPartStride; ++q) acc[q] = part_op(q, acc[q], row_ror<4>(acc[q]));
#pragma unroll
    for (int q = 0; q < kPartStride; ++q) acc[q] = part_op(q, acc[q], row_ror<2>(acc[q]));
#pragma unroll
    for (int q = 0; q < kPartStride; ++q) acc[q] = part_op(q, acc[q], row_ror<1>(acc[q]));
    if ((lane & 15) == 0) {
#pragma unroll
        for (int q = 0; q < kPartStride; ++q) sh[w * 4 + (lane >> 4)][q] = acc[q];
    }
    __syncthreads();
    if (threadIdx.x < kPartStride) {
        const int q = threadIdx.x;
        double a = 0.0;
        for (int k = 0; k < nw; ++k) {
            double r = sh[4 * k][q];
            r = part_op(q, r, sh[4 * k + 1][q]);
            r = part_op(q, r, sh[4 * k + 2][q]);
            r = part_op(q, r, sh[4 * k + 3][q]);
            a = k == 0 ? r : part_op(q, a, r);
        }
        if (q < 5) out[q] = sqrt(a / (double)n);
        else if (q < 10) out[q] = a;
        else if (a > 0.0) *flag = 1;
    }
}

}  // namespace tdt
