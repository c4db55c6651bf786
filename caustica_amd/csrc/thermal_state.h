// Device state of a Pennes/CEM43 session (shared by the thermal kernels and the
// acoustic -> thermal heat-source join).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/caustica_hip.h"
#include "common.h"

struct cst_thermal {
  int device = 0;
  int nx = 0, ny = 0, nz = 0;
  hipStream_t stream = nullptr;
  char* block = nullptr;
  float* theta[2] = {nullptr, nullptr};
  float* tmax = nullptr;
  double* dose = nullptr;
  float* k = nullptr;
  float* perf = nullptr;
  float* coeff = nullptr;
  float* q = nullptr;  // heat source [W/m^3], float32
  unsigned* bad = nullptr;
  int cur = 0;
  int xch2 = 0;  // x planes per block of the two-step kernel; 0: chosen per grid
  // k, perfusion and dt/(rho C) maps all uniform (found at set-up), their values, and
  // whether runs may then take the scalar kernels (cst_thermal_set_uniform_maps)
  bool uniform = false;
  bool allow_uniform = true;
  float ku = 0.f, pfu = 0.f, cfu = 0.f;
  bool have_q = false;
  cst_thermal_desc d{};
  bool ready = false;
  cst::Profiler prof;
};
