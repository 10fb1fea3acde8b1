// tdt_monitor_noalloc: the per-iteration monitor of turbnstwod.f90:147-168
// (L2 and L-infinity of C - COLD per equation, the non-finite flag, then
// COLD <- C) on caller-owned device arrays, with no context and no device
// allocation.  Its kernels are those of tdt_monitor: one partial record per
// block of 256 cells, folded by one block in a fixed order, so the norms have
// the same bits as tdt_monitor's over the same cells in the same order.
//
// The block partials live in `scratch`, a file-scope device array of
// kNoallocMaxCells / 256 records.  Each GPU the library runs on holds its own
// copy until the process ends; every call on a GPU shares it, so calls that
// overlap on one GPU (two streams, two host threads) race.  Every record a
// call folds is written by the same call first, so nothing needs clearing.
#include <mutex>

#include "ctx.hpp"
#include "reduce.hpp"

namespace tdt {
namespace {

constexpr int kNoallocBlocks = TDT_MONITOR_NOALLOC_MAX_CELLS / kBlock;

static __device__ __align__(256) double scratch[(size_t)kNoallocBlocks * kPartStride];

// Device address of `scratch` per device, looked up once per device with that
// device current (the array's name is a host-side stand-in, not an address).
constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

hipError_t scratch_for_current_device(double **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

__global__ void __launch_bounds__(kBlock) k_noalloc_partial(const double *c, long long ldc,
                                                            double *cold, int n, int save_old,
                                                            double *part) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double sq[5] = {0, 0, 0, 0, 0}, mx[5] = {0, 0, 0, 0, 0};
    int bad = 0;
    if (i < n) {
        for (int m = 0; m < 5; ++m) {
            const double cn = c[m * ldc + i];
            const double d = cn - cold[(size_t)m * n + i];
            sq[m] = d * d;
            mx[m] = fabs(d);
            bad |= !isfinite(cn);
            if (save_old) cold[(size_t)m * n + i] = cn;
        }
    }
    block_partials_write(sq, mx, bad, part + (size_t)blockIdx.x * kPartStride);
}

__global__ void __launch_bounds__(kBlock) k_noalloc_fold(const double *part, int nblk, int n,
                                                         double *out) {
    __shared__ int flag;
    if (threadIdx.x == 0) flag = 0;
    __syncthreads();
    reduce_partials(part, nblk, n, out, &flag);
    __syncthreads();
    if (threadIdx.x == 0) out[10] = flag ? 1.0 : 0.0;
}

}  // namespace
}  // namespace tdt

extern "C" int tdt_monitor_noalloc(const double *c, int64_t ldc, double *cold, int32_t n,
                                   int32_t save_old, double *out, void *stream) {
    using namespace tdt;
    if (!c || !cold || !out || n < 1 || ldc < n) {
        api_set_error("tdt_monitor_noalloc: null array, n < 1 or ldc < n");
        return TDT_ERR_ARG;
    }
    if (n > TDT_MONITOR_NOALLOC_MAX_CELLS) {
        api_set_error("tdt_monitor_noalloc: n = " + std::to_string(n) +
                      " exceeds TDT_MONITOR_NOALLOC_MAX_CELLS = " +
                      std::to_string(TDT_MONITOR_NOALLOC_MAX_CELLS));
        return TDT_ERR_ARG;
    }
    double *part = nullptr;
    hipError_t e = scratch_for_current_device(&part);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const int nb = (int)nblocks(n);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_noalloc_partial, dim3(nb), dim3(kBlock), 0, s, c, (long long)ldc,
                           cold, n, save_old ? 1 : 0, part);
        hipLaunchKernelGGL(k_noalloc_fold, dim3(1), dim3(kBlock), 0, s, part, nb, n, out);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        api_set_error(std::string("tdt_monitor_noalloc: ") + hipGetErrorString(e));
        return TDT_ERR_DEVICE;
    }
    return TDT_OK;
}
