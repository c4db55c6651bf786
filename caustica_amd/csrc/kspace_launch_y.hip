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
    dim3 g(strided_blocks(a.s.nx, a.s.nzh, P::TC), ncomp, a.nmem);                           \
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

#define CST_Y3(NAME, KERNEL)                                                          \
  bool NAME(int L, const StepArgs& a, const float2* tw, hipStream_t st) {           \
    switch (L) {                                                                    \
      CST_LENGTHS(CST_Y3_CASE)                                                      \
      default:                                                                      \
        return false;                                                               \
    }                                                                               \
  }

#define CST_Y3_CASE(N)                                                              \
  case N: {                                                                         \
    using P = Plan<N>;                                                              \
    dim3 g(strided_blocks(a.s.nx, a.s.nzh, P::TC), 1, a.nmem);                        \
    hipLaunchKernelGGL((CST_Y3_KERNEL<N>), g, dim3(P::TC * P::T), 0, st, a, tw);     \
    return true;                                                                    \
  }

#define CST_Y3_KERNEL k_ygrad
CST_Y3(launch_ygrad, k_ygrad)
#undef CST_Y3_KERNEL
#define CST_Y3_KERNEL k_ydiv
CST_Y3(launch_ydiv, k_ydiv)
#undef CST_Y3_KERNEL
#undef CST_Y3_CASE
#undef CST_Y3

}  // namespace cst
