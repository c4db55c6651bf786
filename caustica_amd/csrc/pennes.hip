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
// y/z neighbours of plane x come from a double-buffered LDS tile with a one-cell halo.
// Each volume is read once per step (halo re-reads hit L2).
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

// Harmonic mean 2 a b / (a + b) of two conductances with the hardware reciprocal
// (1 ulp) in place of the IEEE division sequence (thirteen instructions, one of them
// quarter rate): the six face means per voxel were two fifths of the kernel's vector
// instructions. Symmetric bitwise in (a, b): 2 a is exact, so 2 a b and 2 b a round
// alike, which lets a plane reuse the x face its predecessor computed.
__device__ __forceinline__ float hmean(float a, float b) {
  return 2.f * a * b * __builtin_amdgcn_rcpf(a + b);
}

// R^(43 - T) with the hardware base-2 exponential: its argument is >= -126 for
// T >= -20 C, where the library exp2f (which only adds a rescaling for arguments below
// -126) returns the same value.
__device__ __forceinline__ float cem43_rate(float t) {
  const float d = t - 43.f;
  return __builtin_amdgcn_exp2f(t > 43.f ? d : 2.f * d);
}

// Software-pipelined march along x: while plane x is computed, the loads for plane x+1
// (its halo and the per-voxel coefficients) and for plane x+2 (centre theta and k) are in
// flight. Two register sets alternate between "current" and "next" (the loop body is
// unrolled twice), so no copy ever waits for a load still in flight. All loads are
// unconditional with indices clamped into the grid (values of the clamped lanes are
// never used), so the wait counts stay exact; the LDS plane is double buffered, one
// barrier per plane.
struct PennesPre {
  float tp, kp;            // centre theta, k of plane x+1 (for plane x)
  float ht, hk;            // theta, k of this thread's halo cell of plane x
  float q, pf, cf, tx;     // heat source, perfusion, dt/(rho C), theta_max of plane x
  double dz;               // dose of plane x
};

// FULL: nx is a multiple of XCH, every block marches exactly XCH planes and the march is
// unrolled completely (no loop header for the wait-count pass to merge at).
template <bool HAS_Q, bool DOSE, bool FULL>
__global__ void __launch_bounds__(TZ* TY) k_pennes(ThermalArgs a) {
  __shared__ float sth[2][TY + 2][TZ + 2];
  __shared__ float sk[2][TY + 2][TZ + 2];
  const int tz = threadIdx.x % TZ, ty = threadIdx.x / TZ;
  const int z = blockIdx.x * TZ + tz;
  const int y = blockIdx.y * TY + ty;
  const int x0 = blockIdx.z * XCH;
  const int x1 = min(x0 + XCH, a.nx);
  const bool valid = (z < a.nz) && (y < a.ny);
  const long plane = (long)a.ny * a.nz;
  const int zc = min(z, a.nz - 1), yc = min(y, a.ny - 1);
  const long col = (long)yc * a.nz + zc;
  // halo: 2 TZ cells of the rows above and below the tile, 2 TY cells before and after
  // it in z. Thread i < 2 TZ loads a row cell (coalesced 256-byte rows), the next 2 TY
  // threads a z cell, the rest none: one load pair per thread instead of a y and a z pair
  // that most lanes would spend re-reading their own cell (clamped at the grid faces:
  // unused there; lanes without a cell read their own and store nothing)
  static_assert(TY * TZ >= 2 * TZ + 2 * TY, "one halo cell per thread at most");
  const int y0 = blockIdx.y * TY, z0 = blockIdx.x * TZ;
  const int hi = threadIdx.x;
  const bool hrow = hi < 2 * TZ, hzc = !hrow && hi < 2 * TZ + 2 * TY;
  const bool hv = hrow || hzc;
  int hly, hlz;  // the cell's place in the LDS tile
  long colh;
  if (hrow) {
    const int side = hi / TZ, zz = min(z0 + hi % TZ, a.nz - 1);
    const int ye = min(y0 + (side ? TY - 1 : 0), a.ny - 1);
    const int yy = side ? min(ye + 1, a.ny - 1) : max(ye - 1, 0);
    hly = side ? TY + 1 : 0;
    hlz = hi % TZ + 1;
    colh = (long)yy * a.nz + zz;
  } else {
    const int j = hzc ? hi - 2 * TZ : 0, row = j >> 1, side = j & 1;
    const int yy = min(y0 + row, a.ny - 1);
    const int ze = min(z0 + (side ? TZ - 1 : 0), a.nz - 1);
    const int zz = side ? min(ze + 1, a.nz - 1) : max(ze - 1, 0);
    hly = row + 1;
    hlz = side ? TZ + 1 : 0;
    colh = hzc ? (long)yy * a.nz + zz : col;
  }
  const int xlast = a.nx - 1;
  auto th = [&](int x, long c) { return a.th_in[(long)x * plane + c]; };
  auto kk = [&](int x, long c) { return a.k[(long)x * plane + c]; };
  // loads for plane x: its halo and coefficients, and the centre of plane x+1
  auto prefetch = [&](int x, PennesPre& p) {
    const int xc = min(x, xlast), xp = min(x + 1, xlast);
    const long i = (long)xc * plane + col;
    p.tp = th(xp, col);
    p.kp = kk(xp, col);
    p.ht = th(xc, colh);
    p.hk = kk(xc, colh);
    p.q = HAS_Q ? a.q[i] : 0.f;
    p.pf = a.perf[i];
    p.cf = a.coeff[i];
    p.tx = a.tmax[i];
    p.dz = DOSE ? a.dose[i] : 0.0;
  };
  const int xm = max(x0 - 1, 0);
  float tm = th(xm, col), km = kk(xm, col);
  float tc = th(x0, col), kc = kk(x0, col);
  float kfm = hmean(km, kc);  // x face below plane x (the previous plane's x face above)
  PennesPre P0, P1;
  prefetch(x0, P0);
  const float iw = a.inv_dx2;
  bool bad = false;
  auto step = [&](int x, int b, const PennesPre& c, PennesPre& n) {
    prefetch(x + 1, n);
    sth[b][ty + 1][tz + 1] = tc;
    sk[b][ty + 1][tz + 1] = kc;
    if (hv) {
      sth[b][hly][hlz] = c.ht;
      sk[b][hly][hlz] = c.hk;
    }
    __syncthreads();
    const float tp = c.tp, kp = c.kp;
    const float kfp = hmean(kc, kp);  // x face above plane x
    if (valid) {
      const long i = (long)x * plane + col;
      float work = 0.f;
      // x faces
      if (x > 0) work += kfm * iw * (tm - tc);
      else if (a.dirichlet & 1) work += 2.f * kc * iw * (a.theta_wall - tc);
      if (x + 1 < a.nx) work += kfp * iw * (tp - tc);
      else if (a.dirichlet & 1) work += 2.f * kc * iw * (a.theta_wall - tc);
      // y faces
      if (y > 0) work += hmean(sk[b][ty][tz + 1], kc) * iw * (sth[b][ty][tz + 1] - tc);
      else if (a.dirichlet & 2) work += 2.f * kc * iw * (a.theta_wall - tc);
      if (y + 1 < a.ny) work += hmean(kc, sk[b][ty + 2][tz + 1]) * iw * (sth[b][ty + 2][tz + 1] - tc);
      else if (a.dirichlet & 2) work += 2.f * kc * iw * (a.theta_wall - tc);
      // z faces
      if (z > 0) work += hmean(sk[b][ty + 1][tz], kc) * iw * (sth[b][ty + 1][tz] - tc);
      else if (a.dirichlet & 4) work += 2.f * kc * iw * (a.theta_wall - tc);
      if (z + 1 < a.nz) work += hmean(kc, sk[b][ty + 1][tz + 2]) * iw * (sth[b][ty + 1][tz + 2] - tc);
      else if (a.dirichlet & 4) work += 2.f * kc * iw * (a.theta_wall - tc);
      if (HAS_Q) work += c.q;
      work -= c.pf * tc;
      const float tn = tc + work * c.cf;
      a.th_out[i] = tn;
      a.tmax[i] = fmaxf(c.tx, tn);
      bad |= !isfinite(tn);
      if (DOSE) {
        const float r_new = cem43_rate(tn + a.t_ref);
        const float r_old = cem43_rate(tc + a.t_ref);
        const double d = c.dz + ((double)r_new + (double)r_old) * a.half_dt_min;
        a.dose[i] = d;
        bad |= !isfinite(d);
      }
    }
    kfm = kfp;
    tm = tc;
    tc = tp;
    kc = kp;
  };
  if constexpr (FULL) {
    static_assert(XCH % 2 == 0, "the unrolled march alternates two register sets");
#pragma unroll
    for (int j = 0; j < XCH; j += 2) {
      step(x0 + j, 0, P0, P1);
      step(x0 + j + 1, 1, P1, P0);
    }
  } else {
    for (int x = x0; x < x1; x += 2) {
      step(x, 0, P0, P1);
      if (x + 1 < x1) step(x + 1, 1, P1, P0);
    }
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
      const bool hq = a.q != nullptr, dd = a.do_dose != 0, full = a.nx % XCH == 0;
      const dim3 bl(TZ * TY);
      hipStream_t st = t->stream;
      if (full) {
        if (hq && dd) hipLaunchKernelGGL((k_pennes<true, true, true>), grid, bl, 0, st, a);
        else if (hq) hipLaunchKernelGGL((k_pennes<true, false, true>), grid, bl, 0, st, a);
        else if (dd) hipLaunchKernelGGL((k_pennes<false, true, true>), grid, bl, 0, st, a);
        else hipLaunchKernelGGL((k_pennes<false, false, true>), grid, bl, 0, st, a);
      } else {
        if (hq && dd) hipLaunchKernelGGL((k_pennes<true, true, false>), grid, bl, 0, st, a);
        else if (hq) hipLaunchKernelGGL((k_pennes<true, false, false>), grid, bl, 0, st, a);
        else if (dd) hipLaunchKernelGGL((k_pennes<false, true, false>), grid, bl, 0, st, a);
        else hipLaunchKernelGGL((k_pennes<false, false, false>), grid, bl, 0, st, a);
      }
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
