// Device check of the hot path's division helpers (csrc/phys.hpp Recip,
// csrc/tmath.hpp tm::quot) against the compiler's IEEE FP64 division, built
// into tests/native/build/libdivcheck.so for tests/test_gpu_division.py.
#include <hip/hip_runtime.h>

#include "phys.hpp"

namespace {

__global__ void k_div(const double *a, const double *b, int n, double *fast, double *ieee,
                      double *quot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = a[i], y = b[i];
    fast[i] = tdt::Recip<double>(y).div(x);
    ieee[i] = x / y;
    quot[i] = tdt::tm::quot(x, y);
}

}  // namespace

// a, b: n host values; fast, ieee, quot: n host results.  Returns 0, or the HIP
// error code.
extern "C" int tdt_divcheck(const double *a, const double *b, int n, double *fast, double *ieee,
                            double *quot) {
    if (n <= 0) return 0;
    const size_t bytes = (size_t)n * sizeof(double);
    double *d = nullptr;
    hipError_t e = hipMalloc(&d, 5 * bytes);
    if (e != hipSuccess) return (int)e;
    double *da = d, *db = d + n, *df = d + 2 * (size_t)n, *di = d + 3 * (size_t)n,
           *dq = d + 4 * (size_t)n;
    e = hipMemcpy(da, a, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_div, dim3((n + 255) / 256), dim3(256), 0, 0, da, db, n, df, di, dq);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(fast, df, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(ieee, di, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(quot, dq, bytes, hipMemcpyDeviceToHost);
    hipError_t e2 = hipFree(d);
    return (int)(e != hipSuccess ? e : e2);
}
