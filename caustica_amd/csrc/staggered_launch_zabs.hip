// Template instantiation + launch of the staggered scheme's pressure z pass of absorbing members.
#include "kspace_launch.h"
#include "staggered_kernels.h"

namespace cst {

template <int N>
static void go_szabs(const StepArgs& a, const float2* tw, const float2* rt, hipStream_t st) {
  using Z = ZPlan<N, 0>;
  const int nl = a.s.nx * a.s.ny;
  hipLaunchKernelGGL((k_szabs<N>), dim3((nl + Z::LPB - 1) / Z::LPB, 1, a.nmem), dim3(Z::THREADS), 0, st,
                     a, tw, rt);
}

bool launch_szabs(int M, const StepArgs& a, const float2* tw, const float2* rt, hipStream_t st) {
  switch (M) {
#define CST_SZA(N)                    \
  case N:                             \
    go_szabs<N>(a, tw, rt, st);       \
    return true;
    CST_LENGTHS(CST_SZA)
#undef CST_SZA
    default:
      return false;
  }
}

}  // namespace cst
