// Template instantiation + launch of the k-space pass kernels: y passes.
#include "kspace_kernels.h"
#include "kspace_launch.h"

namespace cst {

bool launch_y(int L, int dir, const StepArgs& a, const float2* tw, int which, hipStream_t st) {
  const int ncomp = which == 0 ? 1 : 3;
  switch (L) {
#define CST_Y(N)                                                                            \
  case N: {                                                                                 \
    using P = Plan<N>;                                                                      \
    dim3 g(a.s.nx * ((a.s.nzh + P::TC - 1) / P::TC), ncomp, a.nmem);                        \
    if (dir < 0)                                                                            \
      hipLaunchKernelGGL((k_ypass<N, -1>), g, dim3(P::TC * P::T), 0, st, a, tw, which);     \
    else                                                                                    \
      hipLaunchKernelGGL((k_ypass<N, +1>), g, dim3(P::TC * P::T), 0, st, a, tw, which);     \
    return true;                                                                            \
  }
    CST_LENGTHS(CST_Y)
#undef CST_Y
    default:
      return false;
  }
}

}  // namespace cst
