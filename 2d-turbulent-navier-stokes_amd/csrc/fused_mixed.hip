// FP32-mixed instantiation of the fused kernels (fused_kernels.hpp).  This
// translation unit is compiled with -freciprocal-math -fapprox-func
// -ffp-contract=fast: FP32 division and square root become the hardware
// reciprocal / square-root instructions and multiply-adds fuse.  NaN and Inf
// semantics are kept (the non-finite monitor relies on them).
#include "fused_kernels.hpp"

namespace tdt {

hipError_t enqueue_mixed(tdt_ctx *c, FusedBufs &B) {
    if (B.F.ncases) return enqueue_iteration<float, float, float, false, true>(c, B, c->Kf);
    return B.F.has_bc6 ? enqueue_iteration<float, float, float, true, false>(c, B, c->Kf)
                       : enqueue_iteration<float, float, float, false, false>(c, B, c->Kf);
}

}  // namespace tdt
