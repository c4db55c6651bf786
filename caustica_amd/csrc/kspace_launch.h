// Launchers for the k-space pass kernels over the instantiated axis lengths.
#pragma once
#include <hip/hip_runtime.h>

#include "kspace_types.h"

namespace cst {

// Lengths with instantiated kernels: x/y passes use L = nx or ny, z passes use M = nz / 2.
#define CST_LENGTHS(X)                                                                       \
  X(1) X(2) X(4) X(6) X(8) X(10) X(12) X(15) X(16) X(20) X(24) X(30) X(32) X(36) X(40) X(48) \
  X(60) X(64) X(72) X(80) X(96) X(100) X(120) X(128) X(144) X(150) X(160) X(192) X(200)      \
  X(240) X(256) X(288) X(300) X(320) X(384) X(400) X(512)

bool length_supported(int L);

// which: 0 -> sp (one spectrum), 1 -> sg[0..nspec) (each transformed alone);
// tv: tile variant of the strided passes
bool launch_y(int L, int dir, const StepArgs& a, const float2* tw, int which, int tv, hipStream_t st,
              int nspec = 3);
// the 3-spectrum y passes of the step (k_ygrad after xgrad, k_ydiv after zu)
bool launch_ygrad(int L, int tv, const StepArgs& a, const float2* tw, hipStream_t st);
bool launch_ydiv(int L, int tv, const StepArgs& a, const float2* tw, hipStream_t st);
bool launch_xgrad(int L, int tv, const StepArgs& a, const float2* tw, hipStream_t st);
bool launch_xdiv(int L, int tv, const StepArgs& a, const float2* tw, hipStream_t st);
// variant: thread layout and radix plan of the z passes (ZPlan<M, V>), chosen per session
bool launch_zu(int M, int variant, const StepArgs& a, const float2* tw, const float2* rt,
               hipStream_t st);
// velocity component ax between state and real volume (k_zconv): dir > 0 real -> state
bool launch_zconv(int M, const StepArgs& a, const float2* tw, const float2* rt, int ax, int dir,
                  hipStream_t st);
// nonlinear: some member of the batch has a beta map (picks the kernel that loads it)
bool launch_zp(int M, int variant, const StepArgs& a, const float2* tw, const float2* rt, int update,
               bool nonlinear, hipStream_t st);

// ---------------------------------------------------- opt-in staggered scheme
// (staggered_kernels.h; the y passes of that scheme are launch_y's)
bool launch_sxgrad(int L, const StepArgs& a, const float2* tw, hipStream_t st);
bool launch_sygrad(int L, const StepArgs& a, const float2* tw, hipStream_t st);
bool launch_szu(int M, const StepArgs& a, const float2* tw, const float2* rt, hipStream_t st);
bool launch_sydiv(int L, const StepArgs& a, const float2* tw, hipStream_t st);
bool launch_sxdiv(int L, const StepArgs& a, const float2* tw, hipStream_t st);
// absorbing: some member of the batch carries the fractional-Laplacian terms
bool launch_szrho(int M, const StepArgs& a, const float2* tw, const float2* rt, bool absorbing,
                  hipStream_t st);
bool launch_sxabs(int L, const StepArgs& a, const float2* tw, hipStream_t st);
bool launch_szabs(int M, const StepArgs& a, const float2* tw, const float2* rt, hipStream_t st);

}  // namespace cst
