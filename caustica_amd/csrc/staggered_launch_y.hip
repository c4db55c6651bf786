// Template instantiation + launch of the staggered scheme's derivative y passes.
#include "kspace_launch.h"
#include "staggered_kernels.h"

namespace cst {

template <int N>
static void go_sy(bool grad, const StepArgs& a, const float2* tw, hipStream_t st) {
  using P = Plan<N>;
  const unsigned nb = (unsigned)strided_blocks(a.s.nx, a.s.nzh, P::TC);
  if (grad)
    hipLaunchKernelGGL((k_sygrad<N>), dim3(nb, 1, a.nmem), dim3(P::TC * P::T), 0, st, a, tw);
  else
    hipLaunchKernelGGL((k_sydiv<N>), dim3(nb, 3, a.nmem), dim3(P::TC * P::T), 0, st, a, tw);
}

static bool launch_sy(int L, bool grad, const StepArgs& a, const float2* tw, hipStream_t st) {
  switch (L) {
#define CST_SY(N)                   \
  case N:                           \
    go_sy<N>(grad, a, tw, st);      \
    return true;
    CST_LENGTHS(CST_SY)
#undef CST_SY
    default:
      return false;
  }
}

bool launch_sygrad(int L, const StepArgs& a, const float2* tw, hipStream_t st) {
  return launch_sy(L, true, a, tw, st);
}
bool launch_sydiv(int L, const StepArgs& a, const float2* tw, hipStream_t st) {
  return launch_sy(L, false, a, tw, st);
}

}  // namespace cst
