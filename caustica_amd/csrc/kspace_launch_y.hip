// Template instantiation + launch of the k-space pass kernels: y passes.
#include "kspace_kernels.h"
#include "kspace_launch.h"

namespace cst {

template <int N, int TV>
static void go_y(int dir, const StepArgs& a, const float2* tw, int which, int nspec, hipStream_t st) {
  using P = Plan<N, TV>;
  dim3 g(strided_blocks(a.s.nx, a.s.nzh, P::TC), which == 0 ? 1 : nspec, a.nmem);
  if (dir < 0)
    hipLaunchKernelGGL((k_ypass<N, -1, TV>), g, dim3(P::TC * P::T), 0, st, a, tw, which);
  else
    hipLaunchKernelGGL((k_ypass<N, +1, TV>), g, dim3(P::TC * P::T), 0, st, a, tw, which);
}

bool launch_y(int L, int dir, const StepArgs& a, const float2* tw, int which, int tv, hipStream_t st,
              int nspec) {
  if (which != 0 && (nspec < 1 || nspec > 3)) return false;
  switch (L) {
#define CST_Y(N)                                                                           \
  case N:                                                                                  \
    with_tile_variant<N>(tv, [&](auto c) { go_y<N, decltype(c)::value>(dir, a, tw, which, nspec, st); }); \
    return true;
    CST_LENGTHS(CST_Y)
#undef CST_Y
    default:
      return false;
  }
}

template <int N, int TV, bool GRAD>
static void go_y3(const StepArgs& a, const float2* tw, hipStream_t st) {
  using P = Plan<N, TV>;
  dim3 g(strided_blocks(a.s.nx, a.s.nzh, P::TC), 1, a.nmem);
  if (GRAD)
    hipLaunchKernelGGL((k_ygrad<N, TV>), g, dim3(P::TC * P::T), 0, st, a, tw);
  else
    hipLaunchKernelGGL((k_ydiv<N, TV>), g, dim3(P::TC * P::T), 0, st, a, tw);
}

bool launch_ygrad(int L, int tv, const StepArgs& a, const float2* tw, hipStream_t st) {
  switch (L) {
#define CST_YG(N)                                                                          \
  case N:                                                                                  \
    with_tile_variant<N>(tv, [&](auto c) { go_y3<N, decltype(c)::value, true>(a, tw, st); }); \
    return true;
    CST_LENGTHS(CST_YG)
#undef CST_YG
    default:
      return false;
  }
}

bool launch_ydiv(int L, int tv, const StepArgs& a, const float2* tw, hipStream_t st) {
  switch (L) {
#define CST_YD(N)                                                                          \
  case N:                                                                                  \
    with_tile_variant<N>(tv, [&](auto c) { go_y3<N, decltype(c)::value, false>(a, tw, st); }); \
    return true;
    CST_LENGTHS(CST_YD)
#undef CST_YD
    default:
      return false;
  }
}

}  // namespace cst
