// Template instantiation + launch of the staggered scheme's velocity z pass.
#include "kspace_launch.h"
#include "staggered_kernels.h"

namespace cst {

template <int N>
static void go_szu(const StepArgs& a, const float2* tw, const float2* rt, hipStream_t st) {
  using Z = ZPlan<N, 0>;
  const int nl = a.s.nx * a.s.ny;
  const int ng = (nl + Z::LPB - 1) / Z::LPB;
  hipLaunchKernelGGL((k_szu<N>), dim3(24 * ((ng + 7) / 8), 1, a.nmem), dim3(Z::THREADS), 0, st, a, tw, rt);
}

bool launch_szu(int M, const StepArgs& a, const float2* tw, const float2* rt, hipStream_t st) {
  switch (M) {
#define CST_SZU(N)                 \
  case N:                          \
    go_szu<N>(a, tw, rt, st);      \
    return true;
    CST_LENGTHS(CST_SZU)
#undef CST_SZU
    default:
      return false;
  }
}

}  // namespace cst
