// Pennes bioheat + CEM43 dose: one fused explicit step per launch.
//
// theta = T - T_ref (float32), per sub-step (upstream thermal/pennes.py:608-654):
//   work  = sum_faces k_face / dx^2 (theta_nb - theta)   [k_face = 2 k_i k_j / (k_i + k_j)]
//         (+ Dirichlet walls 2 k_edge / dx^2 (theta_wall - theta))
//         + Q - w_b rho_b C_b theta
//   theta' = theta + work * dt_sub / (rho C)
//   theta_max = max(theta_max, theta');  dose += (R^(43-T') + R^(43-T)) * dt_sub / 120
// with R = 0.5 above 43 C and 0.25 at or below (thermal/dose.py:74-89); the dose sum is
// carried in float64.
//
// Blocking: a block owns a TY x TZ tile of the (y, z) plane and marches along x over a
// chunk of XCH planes; theta and k of the planes x-1, x, x+1 live in registers, the
// y/z neighbours of plane x come from an LDS tile with a one-cell halo. Each volume is
// read once per step (halo re-reads hit L2).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/caustica_hip.h"
#include "common.h"
#include "thermal_state.h"

using namespace cst;

namespace {

constexpr int TZ = 64;
constexpr int TY = 4;
constexpr int XCH = 32;

struct ThermalArgs {
  const float* th_in;
  float* th_out;
  float* tmax;
  double* dose;
  const float* k;
  const float* perf;
  const float* coeff;
  const float* q;
  unsigned* bad;
  int nx, ny, nz;
  float inv_dx2;
  int dirichlet;
  float theta_wall;
  float t_ref;
  double half_dt_min;
  int do_dose;
};

__device__ __forceinline__ float hmean(float a, float b) { return 2.f * a * b / (a + b); }

__device__ __forceinline__ float cem43_rate(float t) {
  const float d = t - 43.f;
  return t > 43.f ? exp2f(d) : exp2f(2.f * d);
}

__global__ void __launch_bounds__(TZ* TY) k_pennes(ThermalArgs a) {
  __shared__ float sth[TY + 2][TZ + 2];
  __shared__ float sk[TY + 2][TZ + 2];
  const int tz = threadIdx.x % TZ, ty = threadIdx.x / TZ;
  const int z = blockIdx.x * TZ + tz;
  const int y = blockIdx.y * TY + ty;
  const int x0 = blockIdx.z * XCH;
  const int x1 = min(x0 + XCH, a.nx);
  const bool valid = (z < a.nz) && (y < a.ny);
  const long plane = (long)a.ny * a.nz;
  const long col = valid ? (long)y * a.nz + z : 0;
  auto th = [&](int x) { return a.th_in[(long)x * plane + col]; };
  auto kk = [&](int x) { return a.k[(long)x * plane + col]; };
  float tm = 0.f, km = 1.f, tc = 0.f, kc = 1.f;
  if (valid) {
    tc = th(x0);
    kc = kk(x0);
    if (x0 > 0) {
      tm = th(x0 - 1);
      km = kk(x0 - 1);
    }
  }
  const float iw = a.inv_dx2;
  bool bad = false;
  for (int x = x0; x < x1; ++x) {
    float tp = 0.f, kp = 1.f;
    if (valid && x + 1 < a.nx) {
      tp = th(x + 1);
      kp = kk(x + 1);
    }
    __syncthreads();
    sth[ty + 1][tz + 1] = tc;
    sk[ty + 1][tz + 1] = kc;
    const long pb = (long)x * plane;
    if (ty == 0 && y > 0 && y - 1 < a.ny && z < a.nz) {
      sth[0][tz + 1] = a.th_in[pb + (long)(y - 1) * a.nz + z];
      sk[0][tz + 1] = a.k[pb + (long)(y - 1) * a.nz + z];
    }
    if (ty == TY - 1 && y + 1 < a.ny && z < a.nz) {
      sth[TY + 1][tz + 1] = a.th_in[pb + (long)(y + 1) * a.nz + z];
      sk[TY + 1][tz + 1] = a.k[pb + (long)(y + 1) * a.nz + z];
    }
    if (tz == 0 && z > 0 && y < a.ny) {
      sth[ty + 1][0] = a.th_in[pb + (long)y * a.nz + z - 1];
      sk[ty + 1][0] = a.k[pb + (long)y * a.nz + z - 1];
    }
    if (tz == TZ - 1 && z + 1 < a.nz && y < a.ny) {
      sth[ty + 1][TZ + 1] = a.th_in[pb + (long)y * a.nz + z + 1];
      sk[ty + 1][TZ + 1] = a.k[pb + (long)y * a.nz + z + 1];
    }
    __syncthreads();
    if (valid) {
      const long i = pb + col;
      float work = 0.f;
      // x faces
      if (x > 0) work += hmean(km, kc) * iw * (tm - tc);
      else if (a.dirichlet & 1) work += 2.f * kc * iw * (a.theta_wall - tc);
      if (x + 1 < a.nx) work += hmean(kc, kp) * iw * (tp - tc);
      else if (a.dirichlet & 1) work += 2.f * kc * iw * (a.theta_wall - tc);
      // y faces
      if (y > 0) work += hmean(sk[ty][tz + 1], kc) * iw * (sth[ty][tz + 1] - tc);
      else if (a.dirichlet & 2) work += 2.f * kc * iw * (a.theta_wall - tc);
      if (y + 1 < a.ny) work += hmean(kc, sk[ty + 2][tz + 1]) * iw * (sth[ty + 2][tz + 1] - tc);
      else if (a.dirichlet & 2) work += 2.f * kc * iw * (a.theta_wall - tc);
      // z faces
      if (z > 0) work += hmean(sk[ty + 1][tz], kc) * iw * (sth[ty + 1][tz] - tc);
      else if (a.dirichlet & 4) work += 2.f * kc * iw * (a.theta_wall - tc);
      if (z + 1 < a.nz) work += hmean(kc, sk[ty + 1][tz + 2]) * iw * (sth[ty + 1][tz + 2] - tc);
      else if (a.dirichlet & 4) work += 2.f * kc * iw * (a.theta_wall - tc);
      if (a.q) work += a.q[i];
      work -= a.perf[i] * tc;
      const float tn = tc + work * a.coeff[i];
      a.th_out[i] = tn;
      a.tmax[i] = fmaxf(a.tmax[i], tn);
      bad |= !isfinite(tn);
      if (a.do_dose) {
        const float r_new = cem43_rate(tn + a.t_ref);
        const float r_old = cem43_rate(tc + a.t_ref);
        const double d = a.dose[i] + ((double)r_new + (double)r_old) * a.half_dt_min;
        a.dose[i] = d;
        bad |= !isfinite(d);
      }
    }
    tm = tc;
    km = kc;
    tc = tp;
    kc = kp;
  }
  if (bad) atomicOr(a.bad, 1u);
}

__global__ void k_dose_to_f32(const double* __restrict__ d, float* __restrict__ out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)d[i];
}

__global__ void k_f32_to_dose(const float* __restrict__ in, double* __restrict__ d, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = in ? (double)in[i] : 0.0;
}

}  // namespace

extern "C" {

int cst_thermal_create(int nx, int ny, int nz, int device, cst_thermal** out) {
  if (!out) return fail(CST_ERR_ARG, "null output");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1) return fail(CST_ERR_ARG, "invalid thermal shape");
  HIPCHK(hipSetDevice(device));
  auto* t = new cst_thermal();
  t->device = device;
  t->nx = nx;
  t->ny = ny;
  t->nz = nz;
  const long P = (long)nx * ny * nz;
  Carver c;
  const size_t o_th0 = c.take(P * 4), o_th1 = c.take(P * 4), o_tmax = c.take(P * 4);
  const size_t o_dose = c.take(P * 8), o_k = c.take(P * 4), o_perf = c.take(P * 4);
  const size_t o_coeff = c.take(P * 4), o_q = c.take(P * 4), o_bad = c.take(256);
  hipError_t e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void**)&t->block, c.total);
  if (e != hipSuccess) {
    cst_thermal_destroy(t);
    return fail(CST_ERR_HIP, std::string("thermal allocation: ") + hipGetErrorString(e));
  }
  char* b = t->block;
  t->theta[0] = (float*)(b + o_th0);
  t->theta[1] = (float*)(b + o_th1);
  t->tmax = (float*)(b + o_tmax);
  t->dose = (double*)(b + o_dose);
  t->k = (float*)(b + o_k);
  t->perf = (float*)(b + o_perf);
  t->coeff = (float*)(b + o_coeff);
  t->q = (float*)(b + o_q);
  t->bad = (unsigned*)(b + o_bad);
  *out = t;
  return CST_OK;
}

int cst_thermal_destroy(cst_thermal* t) {
  if (!t) return CST_OK;
  (void)hipSetDevice(t->device);
  if (t->stream) (void)hipStreamSynchronize(t->stream);
  if (t->block) (void)hipFree(t->block);
  t->prof.release();
  if (t->stream) (void)hipStreamDestroy(t->stream);
  delete t;
  return CST_OK;
}

int cst_thermal_setup(cst_thermal* t, const cst_thermal_desc* d) {
  if (!t || !d) return fail(CST_ERR_ARG, "null argument");
  if (!d->k || !d->perf || !d->coeff || !d->theta0) return fail(CST_ERR_ARG, "missing thermal arrays");
  if (!(d->dx > 0)) return fail(CST_ERR_ARG, "dx must be > 0");
  const long P = (long)t->nx * t->ny * t->nz;
  HIPCHK(hipSetDevice(t->device));
  hipStream_t st = t->stream;
  HIPCHK(hipMemcpyAsync(t->k, d->k, P * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(t->perf, d->perf, P * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(t->coeff, d->coeff, P * 4, hipMemcpyHostToDevice, st));
  t->have_q = d->q != nullptr;
  if (d->q) HIPCHK(hipMemcpyAsync(t->q, d->q, P * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(t->theta[0], d->theta0, P * 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(t->tmax, d->theta0, P * 4, hipMemcpyHostToDevice, st));
  // dose0 arrives as float32 (upstream dtype) and is widened on the device
  float* stage = t->theta[1];
  if (d->dose0) HIPCHK(hipMemcpyAsync(stage, d->dose0, P * 4, hipMemcpyHostToDevice, st));
  const int thr = 256;
  const unsigned blocks = (unsigned)((P + thr - 1) / thr);
  hipLaunchKernelGGL(k_f32_to_dose, dim3(blocks), dim3(thr), 0, st, d->dose0 ? stage : nullptr,
                     t->dose, P);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemsetAsync(t->bad, 0, 4, st));
  HIPCHK(hipStreamSynchronize(st));
  t->cur = 0;
  t->d = *d;
  t->ready = true;
  return CST_OK;
}

int cst_thermal_run(cst_thermal* t, int n_substeps) {
  if (!t) return fail(CST_ERR_ARG, "null argument");
  if (!t->ready) return fail(CST_ERR_STATE, "thermal session not set up");
  if (n_substeps < 0) return fail(CST_ERR_ARG, "negative step count");
  HIPCHK(hipSetDevice(t->device));
  ThermalArgs a{};
  a.tmax = t->tmax;
  a.dose = t->dose;
  a.k = t->k;
  a.perf = t->perf;
  a.coeff = t->coeff;
  a.q = t->have_q ? t->q : nullptr;
  a.bad = t->bad;
  a.nx = t->nx;
  a.ny = t->ny;
  a.nz = t->nz;
  a.inv_dx2 = (float)(1.0 / (t->d.dx * t->d.dx));
  a.dirichlet = t->d.boundary;
  a.theta_wall = t->d.theta_wall;
  a.t_ref = t->d.t_ref;
  a.half_dt_min = t->d.half_dt_min;
  a.do_dose = t->d.dose;
  dim3 grid((t->nz + TZ - 1) / TZ, (t->ny + TY - 1) / TY, (t->nx + XCH - 1) / XCH);
  for (int s = 0; s < n_substeps; ++s) {
    a.th_in = t->theta[t->cur];
    a.th_out = t->theta[1 - t->cur];
    t->prof.run(0, t->stream, [&] {
      hipLaunchKernelGGL(k_pennes, grid, dim3(TZ * TY), 0, t->stream, a);
      return true;
    });
    HIPCHK(hipGetLastError());
    t->cur = 1 - t->cur;
  }
  return CST_OK;
}

int cst_thermal_status(cst_thermal* t, int* nonfinite) {
  if (!t || !nonfinite) return fail(CST_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(t->device));
  unsigned v = 0;
  HIPCHK(hipMemcpyAsync(&v, t->bad, 4, hipMemcpyDeviceToHost, t->stream));
  HIPCHK(hipStreamSynchronize(t->stream));
  *nonfinite = v ? 1 : 0;
  return CST_OK;
}

int cst_thermal_read(cst_thermal* t, float* theta, float* theta_max, float* dose) {
  if (!t) return fail(CST_ERR_ARG, "null argument");
  if (!t->ready) return fail(CST_ERR_STATE, "thermal session not set up");
  const long P = (long)t->nx * t->ny * t->nz;
  HIPCHK(hipSetDevice(t->device));
  hipStream_t st = t->stream;
  if (theta) HIPCHK(hipMemcpyAsync(theta, t->theta[t->cur], P * 4, hipMemcpyDeviceToHost, st));
  if (theta_max) HIPCHK(hipMemcpyAsync(theta_max, t->tmax, P * 4, hipMemcpyDeviceToHost, st));
  if (dose) {
    float* stage = t->theta[1 - t->cur];
    const int thr = 256;
    hipLaunchKernelGGL(k_dose_to_f32, dim3((unsigned)((P + thr - 1) / thr)), dim3(thr), 0, st, t->dose,
                       stage, P);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dose, stage, P * 4, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  return CST_OK;
}

int cst_thermal_set_profiling(cst_thermal* t, int on) {
  if (!t) return fail(CST_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(t->device));
  return t->prof.enable(on != 0, t->stream);
}

int cst_thermal_kernel_time(cst_thermal* t, double* total_ms, int64_t* launches) {
  if (!t || !total_ms || !launches) return fail(CST_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(t->device));
  if (int r = t->prof.resolve()) return r;
  *total_ms = t->prof.ms[0];
  *launches = t->prof.count[0];
  return CST_OK;
}

}  // extern "C"
