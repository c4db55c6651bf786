// Template instantiation + launch of the k-space pass kernels: z pressure pass.
#include "kspace_kernels.h"
#include "kspace_launch.h"

namespace cst {

bool launch_zp(int M, const StepArgs& a, const float2* tw, const float2* rt, int update,
               hipStream_t st) {
  const int nl = a.s.nx * a.s.ny;
  switch (M) {
#define CST_ZP(N)                                                                      \
  case N: {                                                                            \
    using Z = ZPlan<N>;                                                                \
    dim3 g((nl + Z::LPB - 1) / Z::LPB, 1, a.nmem);                                     \
    hipLaunchKernelGGL((k_zp<N>), g, dim3(Z::THREADS), 0, st, a, tw, rt, update);      \
    return true;                                                                       \
  }
    CST_LENGTHS(CST_ZP)
#undef CST_ZP
    default:
      return false;
  }
}

}  // namespace cst
