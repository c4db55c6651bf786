// Template instantiation + launch of the k-space pass kernels: x divergence pass.
#include "kspace_kernels.h"
#include "kspace_launch.h"

namespace cst {

template <int N, int TV>
static void go_xdiv(const StepArgs& a, const float2* tw, hipStream_t st) {
  using P = Plan<N, TV>;
  dim3 g((unsigned)strided_blocks(a.s.ny, a.s.nzh, P::TC), 1, a.nmem);
  hipLaunchKernelGGL((k_xdiv<N, TV>), g, dim3(P::TC * P::T), 0, st, a, tw);
}

bool launch_xdiv(int L, int tv, const StepArgs& a, const float2* tw, hipStream_t st) {
  switch (L) {
#define CST_X(N)                                                                   \
  case N:                                                                          \
    with_tile_variant<N>(tv, [&](auto c) { go_xdiv<N, decltype(c)::value>(a, tw, st); }); \
    return true;
    CST_LENGTHS(CST_X)
#undef CST_X
    default:
      return false;
  }
}

}  // namespace cst
