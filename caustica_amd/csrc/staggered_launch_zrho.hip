// Template instantiation + launch of the staggered scheme's density z pass.
#include "kspace_launch.h"
#include "staggered_kernels.h"

namespace cst {

template <int N>
static void go_sz(int kind, const StepArgs& a, const float2* tw, const float2* rt, hipStream_t st) {
  using Z = ZPlan<N, 0>;
  const int nl = a.s.nx * a.s.ny;
  dim3 g((nl + Z::LPB - 1) / Z::LPB, 1, a.nmem);
  if (kind == 0)
    hipLaunchKernelGGL((k_szrho<N, false>), g, dim3(Z::THREADS), 0, st, a, tw, rt);
  else
    hipLaunchKernelGGL((k_szrho<N, true>), g, dim3(Z::THREADS), 0, st, a, tw, rt);
}

static bool launch_sz(int M, int kind, const StepArgs& a, const float2* tw, const float2* rt, hipStream_t st) {
  switch (M) {
#define CST_SZ(N)                     \
  case N:                             \
    go_sz<N>(kind, a, tw, rt, st);    \
    return true;
    CST_LENGTHS(CST_SZ)
#undef CST_SZ
    default:
      return false;
  }
}

bool launch_szrho(int M, const StepArgs& a, const float2* tw, const float2* rt, bool absorbing,
                  hipStream_t st) {
  return launch_sz(M, absorbing ? 1 : 0, a, tw, rt, st);
}

}  // namespace cst
