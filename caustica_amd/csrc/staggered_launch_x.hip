// Template instantiation + launch of the staggered scheme's x passes.
#include "kspace_launch.h"
#include "staggered_kernels.h"

namespace cst {

template <int N>
static void go_sx(int kind, const StepArgs& a, const float2* tw, hipStream_t st) {
  using P = Plan<N>;
  const unsigned nb = (unsigned)strided_blocks(a.s.ny, a.s.nzh, P::TC);
  if (kind == 0)  // the column tiles and, first, the block that advances the step clock
    hipLaunchKernelGGL((k_sxgrad<N>), dim3(nb + 1, 1, a.nmem), dim3(P::TC * P::T), 0, st, a, tw);
  else if (kind == 1)
    hipLaunchKernelGGL((k_sxdiv<N>), dim3(nb, 3, a.nmem), dim3(P::TC * P::T), 0, st, a, tw);
  else
    hipLaunchKernelGGL((k_sxabs<N>), dim3(nb, 2, a.nmem), dim3(P::TC * P::T), 0, st, a, tw);
}

static bool launch_sx(int L, int kind, const StepArgs& a, const float2* tw, hipStream_t st) {
  switch (L) {
#define CST_SX(N)                    \
  case N:                            \
    go_sx<N>(kind, a, tw, st);       \
    return true;
    CST_LENGTHS(CST_SX)
#undef CST_SX
    default:
      return false;
  }
}

bool launch_sxgrad(int L, const StepArgs& a, const float2* tw, hipStream_t st) {
  return launch_sx(L, 0, a, tw, st);
}
bool launch_sxdiv(int L, const StepArgs& a, const float2* tw, hipStream_t st) {
  return launch_sx(L, 1, a, tw, st);
}
bool launch_sxabs(int L, const StepArgs& a, const float2* tw, hipStream_t st) {
  return launch_sx(L, 2, a, tw, st);
}

}  // namespace cst
