// Template instantiation + launch of the k-space pass kernels: z pressure pass.
#include "kspace_kernels.h"
#include "kspace_launch.h"

namespace cst {

template <int N, int V>
static void go_zp(const StepArgs& a, const float2* tw, const float2* rt, int update, bool nonlinear,
                  hipStream_t st) {
  using Z = ZPlan<N, V>;
  const int nl = a.s.nx * a.s.ny;
  dim3 g((nl + Z::LPB - 1) / Z::LPB, 1, a.nmem);
  if (nonlinear)
    hipLaunchKernelGGL((k_zp<N, V, true>), g, dim3(Z::THREADS), 0, st, a, tw, rt, update);
  else
    hipLaunchKernelGGL((k_zp<N, V, false>), g, dim3(Z::THREADS), 0, st, a, tw, rt, update);
}

bool launch_zp(int M, int variant, const StepArgs& a, const float2* tw, const float2* rt, int update,
               bool nonlinear, hipStream_t st) {
  switch (M) {
#define CST_ZP(N)                                                                    \
  case N:                                                                            \
    with_z_variant<N, 2>(variant, [&](auto c) {                                      \
      go_zp<N, decltype(c)::value>(a, tw, rt, update, nonlinear, st);                \
    });                                                                              \
    return true;
    CST_LENGTHS(CST_ZP)
#undef CST_ZP
    default:
      return false;
  }
}

}  // namespace cst
