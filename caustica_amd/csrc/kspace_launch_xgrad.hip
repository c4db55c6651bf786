// Template instantiation + launch of the k-space pass kernels: x gradient pass.
#include "kspace_kernels.h"
#include "kspace_launch.h"

namespace cst {

bool launch_xgrad(int L, const StepArgs& a, const float2* tw, hipStream_t st) {
  switch (L) {
#define CST_XG(N)                                                                \
  case N: {                                                                      \
    using P = Plan<N>;                                                           \
    dim3 g((unsigned)strided_blocks(a.s.ny, a.s.nzh, P::TC), 1, a.nmem);         \
    hipLaunchKernelGGL((k_xgrad<N>), g, dim3(P::TC * P::T), 0, st, a, tw);       \
    return true;                                                                 \
  }
    CST_LENGTHS(CST_XG)
#undef CST_XG
    default:
      return false;
  }
}

}  // namespace cst
