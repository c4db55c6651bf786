// Template instantiation + launch of the k-space pass kernels: x gradient pass.
#include "kspace_kernels.h"
#include "kspace_launch.h"

namespace cst {

template <int N, int TV>
static void go_xgrad(const StepArgs& a, const float2* tw, hipStream_t st) {
  using P = Plan<N, TV>;
  // the column tiles and, last, the block that advances the step clock
  dim3 g((unsigned)strided_blocks(a.s.ny, a.s.nzh, P::TC) + 1, 1, a.nmem);
  hipLaunchKernelGGL((k_xgrad<N, TV>), g, dim3(P::TC * P::T), 0, st, a, tw);
}

bool launch_xgrad(int L, int tv, const StepArgs& a, const float2* tw, hipStream_t st) {
  switch (L) {
#define CST_X(N)                                                                   \
  case N:                                                                          \
    with_tile_variant<N>(tv, [&](auto c) { go_xgrad<N, decltype(c)::value>(a, tw, st); }); \
    return true;
    CST_LENGTHS(CST_X)
#undef CST_X
    default:
      return false;
  }
}

}  // namespace cst
