// Template instantiation + launch of the k-space pass kernels: z velocity pass.
#include "kspace_kernels.h"
#include "kspace_launch.h"

namespace cst {

template <int N, int V>
static void go_zu(const StepArgs& a, const float2* tw, const float2* rt, hipStream_t st) {
  using Z = ZPlan<N, V>;
  const int nl = a.s.nx * a.s.ny;
  const int ng = (nl + Z::LPB - 1) / Z::LPB;
  dim3 g(24 * ((ng + 7) / 8), 1, a.nmem);
  hipLaunchKernelGGL((k_zu<N, V>), g, dim3(Z::THREADS), 0, st, a, tw, rt);
}

bool launch_zu(int M, int variant, const StepArgs& a, const float2* tw, const float2* rt,
               hipStream_t st) {
  switch (M) {
#define CST_ZU(N)                                                                   \
  case N:                                                                           \
    with_z_variant<N, 3>(variant, [&](auto c) { go_zu<N, decltype(c)::value>(a, tw, rt, st); }); \
    return true;
    CST_LENGTHS(CST_ZU)
#undef CST_ZU
    default:
      return false;
  }
}

template <int N>
static void go_zconv(const StepArgs& a, const float2* tw, const float2* rt, int ax, int dir, hipStream_t st) {
  using Z = ZPlan<N, 0>;
  const int nl = a.s.nx * a.s.ny;
  dim3 g((nl + Z::LPB - 1) / Z::LPB, 1, a.nmem);
  hipLaunchKernelGGL((k_zconv<N, 0>), g, dim3(Z::THREADS), 0, st, a, tw, rt, ax, dir);
}

bool launch_zconv(int M, const StepArgs& a, const float2* tw, const float2* rt, int ax, int dir,
                  hipStream_t st) {
  switch (M) {
#define CST_ZC(N)                              \
  case N:                                      \
    go_zconv<N>(a, tw, rt, ax, dir, st);       \
    return true;
    CST_LENGTHS(CST_ZC)
#undef CST_ZC
    default:
      return false;
  }
}

}  // namespace cst
