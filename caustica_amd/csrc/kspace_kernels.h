// Pass kernels of the CW k-space step (one template instantiation per axis length).
//
// One acoustic step is eight launches over the half-spectrum layout (see kspace_types.h):
//   xgrad  x-FFT of p-hat, times i k_a kappa / N for a = x, y, z, inverse x-FFT -> sg[a]
//   y inv  (3 spectra)
//   zu     z-C2R of each sg[a] -> grad_a; u_a = ((u_a - dt/rho grad_a) absorb) sponge;
//          z-R2C of u_a -> sg[a]
//   y fwd  (3 spectra)
//   xdiv   x-FFT of each sg[a], sum_a i k_a kappa / N * U_a, inverse x-FFT -> sp
//   y inv  (1 spectrum)
//   zp     z-C2R of sp -> div u; p update (+ Westervelt), absorb, sponge, source, peak,
//          record; z-R2C of p -> sp
//   y fwd  (1 spectrum)
// Every line FFT is a two-stage Stockham-style split L = A * B across T threads of a
// line, with the exchange and the twiddle table in LDS (see Plan<L>).
#pragma once
#include <utility>

#include "fft_codelets.h"
#include "kspace_types.h"

namespace cst {

constexpr int plan_a(int L) {
  for (int a = 1; a <= L; ++a)
    if (L % a == 0 && a * a >= L) return a;
  return L;
}
constexpr int cgcd(int a, int b) { return b == 0 ? a : cgcd(b, a % b); }
constexpr int pow2_floor(int v) {
  int p = 1;
  while (p * 2 <= v) p *= 2;
  return p;
}
constexpr int cmin(int a, int b) { return a < b ? a : b; }

// L = A * B; step 1: A DFTs of size B; step 2: B DFTs of size A; T threads per line.
template <int L>
struct Plan {
  static constexpr int A = plan_a(L);
  static constexpr int B = L / A;
  static constexpr int G = cgcd(A, B);
  static constexpr int T = G >= 4 ? G : B;
  static constexpr int R1 = (A + T - 1) / T;
  static constexpr int R2 = (B + T - 1) / T;
  // strided passes: columns per block (tile of L x TC complex in LDS, <= 48 KiB)
  static constexpr int TC = pow2_floor(cmin(256 / T, 6144 / L)) < 1 ? 1 : pow2_floor(cmin(256 / T, 6144 / L));
};

// z passes: complex length M = nz / 2, LPB lines per block
template <int M>
struct ZPlan {
  using P = Plan<M>;
  static constexpr int LS = M + 2;  // per-line LDS stride (float2), holds M + 1 spectrum bins
  static constexpr int LPB_raw = pow2_floor(cmin(256 / P::T, 5120 / LS));
  static constexpr int LPB = LPB_raw < 1 ? 1 : LPB_raw;
  static constexpr int THREADS = LPB * P::T;
};

template <int DIR>
__device__ __forceinline__ float2 twmul(float2 v, float2 w) {
  return DIR > 0 ? cmul(v, w) : cmulconj(v, w);
}

// ---------------------------------------------------------------- stages
// stage "in": natural-order input x[n1 + A n2]  ->  X[k2 + B k1] held as X[r][k1], k2 = t + T r
template <int L, class Load>
__device__ __forceinline__ void in_load(int t, Load load, float2 (&V)[Plan<L>::R1][Plan<L>::B]) {
  using P = Plan<L>;
#pragma unroll
  for (int r = 0; r < P::R1; ++r) {
    const int n1 = t + P::T * r;
#pragma unroll
    for (int n2 = 0; n2 < P::B; ++n2) V[r][n2] = (n1 < P::A) ? load(n1 + P::A * n2) : make_float2(0.f, 0.f);
  }
}

template <int L, int DIR, class At>
__device__ __forceinline__ void in_compute(int t, float2 (&V)[Plan<L>::R1][Plan<L>::B], At at,
                                           const float2* tw) {
  using P = Plan<L>;
#pragma unroll
  for (int r = 0; r < P::R1; ++r) {
    const int n1 = t + P::T * r;
    if (n1 < P::A) {
      Dft<P::B, DIR>::run(V[r]);
#pragma unroll
      for (int k2 = 0; k2 < P::B; ++k2) *at(n1 + P::A * k2) = twmul<DIR>(V[r][k2], tw[n1 * k2]);
    }
  }
}

template <int L, int DIR, class At>
__device__ __forceinline__ void in_finish(int t, At at, float2 (&X)[Plan<L>::R2][Plan<L>::A]) {
  using P = Plan<L>;
#pragma unroll
  for (int r = 0; r < P::R2; ++r) {
    const int k2 = t + P::T * r;
    if (k2 < P::B) {
#pragma unroll
      for (int n1 = 0; n1 < P::A; ++n1) X[r][n1] = *at(n1 + P::A * k2);
      Dft<P::A, DIR>::run(X[r]);
    }
  }
}

// stage "out": input X[k2 + B k1] as X[r][k1]  ->  natural x[m2 + A j] as V[r][j], m2 = t + T r
template <int L, int DIR, class At>
__device__ __forceinline__ void out_compute(int t, float2 (&X)[Plan<L>::R2][Plan<L>::A], At at,
                                            const float2* tw) {
  using P = Plan<L>;
#pragma unroll
  for (int r = 0; r < P::R2; ++r) {
    const int k2 = t + P::T * r;
    if (k2 < P::B) {
      float2 v[P::A];
#pragma unroll
      for (int i = 0; i < P::A; ++i) v[i] = X[r][i];
      Dft<P::A, DIR>::run(v);
#pragma unroll
      for (int m2 = 0; m2 < P::A; ++m2) *at(k2 + P::B * m2) = twmul<DIR>(v[m2], tw[k2 * m2]);
    }
  }
}

template <int L, int DIR, class At>
__device__ __forceinline__ void out_finish(int t, At at, float2 (&V)[Plan<L>::R1][Plan<L>::B]) {
  using P = Plan<L>;
#pragma unroll
  for (int r = 0; r < P::R1; ++r) {
    const int m2 = t + P::T * r;
    if (m2 < P::A) {
#pragma unroll
      for (int k2 = 0; k2 < P::B; ++k2) V[r][k2] = *at(k2 + P::B * m2);
      Dft<P::B, DIR>::run(V[r]);
    }
  }
}

template <int TCP>
struct TileAt {
  float2* base;
  int col;
  __device__ __forceinline__ float2* operator()(int row) const { return base + row * TCP + col; }
};

struct LineAt {
  float2* base;
  __device__ __forceinline__ float2* operator()(int row) const { return base + row; }
};

__device__ __forceinline__ float sincf_half(float x) {  // sin(x) / x
  return x == 0.f ? 1.f : sinf(x) / x;
}

// ================================================================ y pass
// In-place C2C along y of one or three half spectra. Columns c = x * nzh + zb.
template <int L, int DIR>
__global__ void __launch_bounds__(Plan<L>::TC* Plan<L>::T)
    k_ypass(StepArgs a, const float2* __restrict__ twg, int which) {
  using P = Plan<L>;
  constexpr int TC = P::TC;
  __shared__ float2 tile[L * TC];
  __shared__ float2 tw[L];
  for (int i = threadIdx.x; i < L; i += TC * P::T) tw[i] = twg[i];
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  float2* __restrict__ data = which == 0 ? m.sp : m.sg[blockIdx.y];
  const int col = threadIdx.x % TC, t = threadIdx.x / TC;
  const int nzh = a.s.nzh;
  const int C = a.s.nx * nzh;
  const int c = blockIdx.x * TC + col;
  const bool valid = c < C;
  const int cc = valid ? c : 0;
  const long base = (long)(cc / nzh) * L * nzh + (cc % nzh);
  const long S = nzh;
  __syncthreads();
  float2 V[P::R1][P::B];
  in_load<L>(t, [&](int n) { return valid ? data[base + n * S] : make_float2(0.f, 0.f); }, V);
  TileAt<TC> at{tile, col};
  in_compute<L, DIR>(t, V, at, tw);
  __syncthreads();
  float2 X[P::R2][P::A];
  in_finish<L, DIR>(t, at, X);
  if (valid) {
#pragma unroll
    for (int r = 0; r < P::R2; ++r) {
      const int k2 = t + P::T * r;
      if (k2 < P::B) {
#pragma unroll
        for (int k1 = 0; k1 < P::A; ++k1) data[base + (long)(k2 + P::B * k1) * S] = X[r][k1];
      }
    }
  }
}

// ================================================================ x passes
// Columns c = y * nzh + zb are contiguous; element (kx, c) at kx * ny * nzh + c.
template <int L>
__device__ __forceinline__ void x_kappa(const MemberDev& m, const Shape& s, int c, int t,
                                        float (&kap)[Plan<L>::R2][Plan<L>::A], float& kyd,
                                        float& kzd) {
  using P = Plan<L>;
  const float* kt = m.ktab;
  const float* kx_full = kt;
  const float* ky_full = kt + 2 * s.nx;
  const float* ky_d = ky_full + s.ny;
  const float* kz_full = ky_d + s.ny;
  const float* kz_d = kz_full + s.nzh;
  const int y = c / s.nzh, zb = c - y * s.nzh;
  const float kyf = ky_full[y], kzf = kz_full[zb];
  const float kyz2 = kyf * kyf + kzf * kzf;
  kyd = ky_d[y];
  kzd = kz_d[zb];
#pragma unroll
  for (int r = 0; r < P::R2; ++r) {
    const int k2 = t + P::T * r;
#pragma unroll
    for (int k1 = 0; k1 < P::A; ++k1) {
      const int kxi = (k2 < P::B) ? k2 + P::B * k1 : 0;
      const float kxf = kx_full[kxi];
      kap[r][k1] = sincf_half(m.half_cdt * sqrtf(kxf * kxf + kyz2)) * m.inv_n;
    }
  }
}

// gradient: sg[a] = IFFT_x( i k_a kappa / N * FFT_x(sp) ),  a = x, y, z
template <int L>
__global__ void __launch_bounds__(Plan<L>::TC* Plan<L>::T)
    k_xgrad(StepArgs a, const float2* __restrict__ twg) {
  using P = Plan<L>;
  constexpr int TC = P::TC;
  __shared__ float2 tile[L * TC];
  __shared__ float2 tw[L];
  for (int i = threadIdx.x; i < L; i += TC * P::T) tw[i] = twg[i];
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  const int col = threadIdx.x % TC, t = threadIdx.x / TC;
  const long C = (long)a.s.ny * a.s.nzh;
  const long c = (long)blockIdx.x * TC + col;
  const bool valid = c < C;
  const long base = valid ? c : 0;
  __syncthreads();
  float2 V[P::R1][P::B];
  const float2* __restrict__ src = m.sp;
  in_load<L>(t, [&](int n) { return valid ? src[base + n * C] : make_float2(0.f, 0.f); }, V);
  TileAt<TC> at{tile, col};
  in_compute<L, -1>(t, V, at, tw);
  __syncthreads();
  float2 X[P::R2][P::A];
  in_finish<L, -1>(t, at, X);
  float kap[P::R2][P::A];
  float kyd, kzd;
  x_kappa<L>(m, a.s, (int)base, t, kap, kyd, kzd);
  const float* kx_d = m.ktab + a.s.nx;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    float2 Z[P::R2][P::A];
#pragma unroll
    for (int r = 0; r < P::R2; ++r) {
      const int k2 = t + P::T * r;
#pragma unroll
      for (int k1 = 0; k1 < P::A; ++k1) {
        const int kxi = (k2 < P::B) ? k2 + P::B * k1 : 0;
        const float kd = ax == 0 ? kx_d[kxi] : (ax == 1 ? kyd : kzd);
        const float d = kd * kap[r][k1];
        Z[r][k1] = make_float2(-d * X[r][k1].y, d * X[r][k1].x);
      }
    }
    __syncthreads();
    out_compute<L, +1>(t, Z, at, tw);
    __syncthreads();
    out_finish<L, +1>(t, at, V);
    if (valid) {
      float2* __restrict__ dst = m.sg[ax];
#pragma unroll
      for (int r = 0; r < P::R1; ++r) {
        const int m2 = t + P::T * r;
        if (m2 < P::A) {
#pragma unroll
          for (int j = 0; j < P::B; ++j) dst[base + (long)(m2 + P::A * j) * C] = V[r][j];
        }
      }
    }
  }
}

// divergence: sp = IFFT_x( sum_a i k_a kappa / N * FFT_x(sg[a]) )
template <int L>
__global__ void __launch_bounds__(Plan<L>::TC* Plan<L>::T)
    k_xdiv(StepArgs a, const float2* __restrict__ twg) {
  using P = Plan<L>;
  constexpr int TC = P::TC;
  __shared__ float2 tile[L * TC];
  __shared__ float2 tw[L];
  for (int i = threadIdx.x; i < L; i += TC * P::T) tw[i] = twg[i];
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  const int col = threadIdx.x % TC, t = threadIdx.x / TC;
  const long C = (long)a.s.ny * a.s.nzh;
  const long c = (long)blockIdx.x * TC + col;
  const bool valid = c < C;
  const long base = valid ? c : 0;
  float kap[P::R2][P::A];
  float kyd, kzd;
  x_kappa<L>(m, a.s, (int)base, t, kap, kyd, kzd);
  const float* kx_d = m.ktab + a.s.nx;
  TileAt<TC> at{tile, col};
  float2 Sacc[P::R2][P::A];
  float2 V[P::R1][P::B];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const float2* __restrict__ src = m.sg[ax];
    in_load<L>(t, [&](int n) { return valid ? src[base + n * C] : make_float2(0.f, 0.f); }, V);
    __syncthreads();
    in_compute<L, -1>(t, V, at, tw);
    __syncthreads();
    float2 X[P::R2][P::A];
    in_finish<L, -1>(t, at, X);
#pragma unroll
    for (int r = 0; r < P::R2; ++r) {
      const int k2 = t + P::T * r;
#pragma unroll
      for (int k1 = 0; k1 < P::A; ++k1) {
        const int kxi = (k2 < P::B) ? k2 + P::B * k1 : 0;
        const float kd = ax == 0 ? kx_d[kxi] : (ax == 1 ? kyd : kzd);
        const float d = kd * kap[r][k1];
        const float2 term = make_float2(-d * X[r][k1].y, d * X[r][k1].x);
        Sacc[r][k1] = ax == 0 ? term : cadd(Sacc[r][k1], term);
      }
    }
  }
  __syncthreads();
  out_compute<L, +1>(t, Sacc, at, tw);
  __syncthreads();
  out_finish<L, +1>(t, at, V);
  if (valid) {
    float2* __restrict__ dst = m.sp;
#pragma unroll
    for (int r = 0; r < P::R1; ++r) {
      const int m2 = t + P::T * r;
      if (m2 < P::A) {
#pragma unroll
        for (int j = 0; j < P::B; ++j) dst[base + (long)(m2 + P::A * j) * C] = V[r][j];
      }
    }
  }
}

// ================================================================ z passes
// Real length nz = 2M along the contiguous axis, computed as a length-M complex FFT of
// z[m] = x[2m] + i x[2m+1] with the even/odd split in LDS (rt[k] = exp(2 pi i k / nz)).

// Load M+1 half-spectrum bins of the block's lines into LDS (block-contiguous in HBM).
template <int M>
__device__ __forceinline__ void z_stage_spectrum(const float2* __restrict__ spec, long line0,
                                                 int nlines_total, float2* lb) {
  using Z = ZPlan<M>;
  constexpr int NB = M + 1;
  const long off = line0 * NB;
  const int nl = cmin(Z::LPB, nlines_total - (int)line0);
  for (int i = threadIdx.x; i < nl * NB; i += Z::THREADS) {
    const int l = i / NB, k = i - l * NB;
    lb[l * Z::LS + k] = spec[off + i];
  }
}

// C2R pre-split: Z[k] = Fe[k] + i Fo[k], Fe = X[k] + conj(X[M-k]),
// Fo = (X[k] - conj(X[M-k])) exp(+2 pi i k / nz); imaginary parts of X[0], X[M] ignored.
template <int M>
__device__ __forceinline__ float2 c2r_pre(const float2* l, int k, const float2* rt) {
  float2 xk = l[k], xm = l[M - k];
  if (k == 0) {
    xk.y = 0.f;
    xm.y = 0.f;
  }
  const float2 fe = make_float2(xk.x + xm.x, xk.y - xm.y);
  const float2 d = make_float2(xk.x - xm.x, xk.y + xm.y);
  const float2 fo = cmul(d, rt[k]);
  return make_float2(fe.x - fo.y, fe.y + fo.x);
}

// R2C post-split from natural Z[0..M) in LDS: X[k] = E + exp(-2 pi i k / nz) O,
// E = (Z[k] + conj(Z[M-k])) / 2, O = -i (Z[k] - conj(Z[M-k])) / 2.
template <int M>
__device__ __forceinline__ float2 r2c_post(const float2* l, int k, const float2* rt) {
  const float2 zk = l[k == M ? 0 : k];
  const float2 zm = l[k == 0 ? 0 : M - k];
  const float2 e = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
  const float2 o = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
  const float2 w = cmulconj(o, rt[k]);
  float2 x = cadd(e, w);
  if (k == 0 || k == M) x.y = 0.f;
  return x;
}

// Forward R2C of X (thread-held, z[m] = x[2m] + i x[2m+1], m = k2 + B k1) to HBM.
template <int M>
__device__ __forceinline__ void z_r2c_store(int t, float2 (&X)[Plan<M>::R2][Plan<M>::A], float2* lb,
                                            float2* lb_line, const float2* tw, const float2* rt,
                                            float2* __restrict__ dst_line, bool valid) {
  using P = Plan<M>;
  LineAt at{lb_line};
  __syncthreads();
  out_compute<M, -1>(t, X, at, tw);
  __syncthreads();
  float2 V[P::R1][P::B];
  out_finish<M, -1>(t, at, V);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < P::R1; ++r) {
    const int m2 = t + P::T * r;
    if (m2 < P::A) {
#pragma unroll
      for (int j = 0; j < P::B; ++j) lb_line[m2 + P::A * j] = V[r][j];
    }
  }
  __syncthreads();
  if (valid) {
    for (int k = t; k <= M; k += P::T) dst_line[k] = r2c_post<M>(lb_line, k, rt);
  }
}

// Inverse C2R of the spectrum line staged in LDS into thread-held z pairs X[r][k1].
template <int M>
__device__ __forceinline__ void z_c2r_regs(int t, float2* lb_line, const float2* tw, const float2* rt,
                                           float2 (&X)[Plan<M>::R2][Plan<M>::A]) {
  using P = Plan<M>;
  LineAt at{lb_line};
  float2 V[P::R1][P::B];
  in_load<M>(t, [&](int n) { return c2r_pre<M>(lb_line, n, rt); }, V);
  __syncthreads();
  in_compute<M, +1>(t, V, at, tw);
  __syncthreads();
  in_finish<M, +1>(t, at, X);
}

template <int M>
__device__ __forceinline__ void z_tables(const float2* twg, const float2* rtg, float2* tw, float2* rt) {
  using Z = ZPlan<M>;
  for (int i = threadIdx.x; i < M; i += Z::THREADS) tw[i] = twg[i];
  for (int i = threadIdx.x; i <= M; i += Z::THREADS) rt[i] = rtg[i];
}

// velocity update for the three axes
template <int M>
__global__ void __launch_bounds__(ZPlan<M>::THREADS)
    k_zu(StepArgs a, const float2* __restrict__ twg, const float2* __restrict__ rtg) {
  using P = Plan<M>;
  using Z = ZPlan<M>;
  __shared__ float2 lb[Z::LPB * Z::LS];
  __shared__ float2 tw[M];
  __shared__ float2 rt[M + 1];
  z_tables<M>(twg, rtg, tw, rt);
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  const int nlines = a.s.nx * a.s.ny;
  const int nz = 2 * M;
  const long line0 = (long)blockIdx.x * Z::LPB;
  const int ll = threadIdx.x / P::T, t = threadIdx.x % P::T;
  const long line = line0 + ll;
  const bool valid = line < nlines;
  const long lsafe = valid ? line : 0;
  float2* lbl = lb + ll * Z::LS;
  const int x = (int)(lsafe / a.s.ny), y = (int)(lsafe - (long)x * a.s.ny);
  const float* sp = m.sponge;
  const float sxy = sp[nz + x] * sp[nz + a.s.nx + y];
  const float* sz = sp;
  const float2* __restrict__ dto = reinterpret_cast<const float2*>(m.dto_rho + lsafe * nz);
  const float2* __restrict__ ab = reinterpret_cast<const float2*>(m.absorb + lsafe * nz);
#pragma unroll 1
  for (int ax = 0; ax < 3; ++ax) {
    float2* __restrict__ spec = m.sg[ax];
    __syncthreads();
    z_stage_spectrum<M>(spec, line0, nlines, lb);
    __syncthreads();
    float2 X[P::R2][P::A];
    z_c2r_regs<M>(t, lbl, tw, rt, X);
    float2* __restrict__ u = reinterpret_cast<float2*>(m.u[ax] + lsafe * nz);
#pragma unroll
    for (int r = 0; r < P::R2; ++r) {
      const int k2 = t + P::T * r;
      if (k2 < P::B) {
#pragma unroll
        for (int k1 = 0; k1 < P::A; ++k1) {
          const int mm = k2 + P::B * k1;
          if (valid) {
            const float2 g = X[r][k1];
            float2 uv = u[mm];
            const float2 d = dto[mm];
            const float2 b = ab[mm];
            const float2 s2 = reinterpret_cast<const float2*>(sz)[mm];
            uv.x -= d.x * g.x;
            uv.y -= d.y * g.y;
            uv.x *= b.x;
            uv.y *= b.y;
            uv.x *= sxy * s2.x;
            uv.y *= sxy * s2.y;
            u[mm] = uv;
            X[r][k1] = uv;
          }
        }
      }
    }
    z_r2c_store<M>(t, X, lb, lbl, tw, rt, spec + lsafe * (M + 1), valid);
  }
}

// pressure update (+ Westervelt), source, convergence peak, record; or (mode 0) a plain
// R2C of the current p.
template <int M>
__global__ void __launch_bounds__(ZPlan<M>::THREADS)
    k_zp(StepArgs a, const float2* __restrict__ twg, const float2* __restrict__ rtg, int update) {
  using P = Plan<M>;
  using Z = ZPlan<M>;
  __shared__ float2 lb[Z::LPB * Z::LS];
  __shared__ float2 tw[M];
  __shared__ float2 rt[M + 1];
  z_tables<M>(twg, rtg, tw, rt);
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  const StepMember& sm = a.sm[blockIdx.z];
  const int nlines = a.s.nx * a.s.ny;
  const int nz = 2 * M;
  const long line0 = (long)blockIdx.x * Z::LPB;
  const int ll = threadIdx.x / P::T, t = threadIdx.x % P::T;
  const long line = line0 + ll;
  const bool valid = line < nlines;
  const long lsafe = valid ? line : 0;
  float2* lbl = lb + ll * Z::LS;
  const int x = (int)(lsafe / a.s.ny), y = (int)(lsafe - (long)x * a.s.ny);
  float2* __restrict__ pl = reinterpret_cast<float2*>(m.p + lsafe * nz);
  float2 X[P::R2][P::A];
  if (update) {
    __syncthreads();
    z_stage_spectrum<M>(m.sp, line0, nlines, lb);
    __syncthreads();
    z_c2r_regs<M>(t, lbl, tw, rt, X);
    const float* sp = m.sponge;
    const float sxy = sp[nz + x] * sp[nz + a.s.nx + y];
    const float2* __restrict__ sz = reinterpret_cast<const float2*>(sp);
    const float2* __restrict__ rc = reinterpret_cast<const float2*>(m.rhoc2dt + lsafe * nz);
    const float2* __restrict__ ab = reinterpret_cast<const float2*>(m.absorb + lsafe * nz);
    const float2* __restrict__ bt =
        m.beta2dt ? reinterpret_cast<const float2*>(m.beta2dt + lsafe * nz) : nullptr;
#pragma unroll
    for (int r = 0; r < P::R2; ++r) {
      const int k2 = t + P::T * r;
      if (k2 < P::B) {
#pragma unroll
        for (int k1 = 0; k1 < P::A; ++k1) {
          const int mm = k2 + P::B * k1;
          if (valid) {
            const float2 dv = X[r][k1];
            float2 pv = pl[mm];
            float2 c = rc[mm];
            if (bt) {
              const float2 b = bt[mm];
              c.x += b.x * pv.x;
              c.y += b.y * pv.y;
            }
            const float2 b = ab[mm];
            const float2 s2 = sz[mm];
            pv.x -= c.x * dv.x;
            pv.y -= c.y * dv.y;
            pv.x *= b.x;
            pv.y *= b.y;
            pv.x *= sxy * s2.x;
            pv.y *= sxy * s2.y;
            X[r][k1] = pv;
          }
        }
      }
    }
    // source voxels on this line
    if (valid) {
      const int beg = m.src_ptr[lsafe], end = m.src_ptr[lsafe + 1];
      for (int j = beg; j < end; ++j) {
        const int zz = m.src_z[j];
        const int mm = zz >> 1;
        const int k2 = mm % P::B, k1s = mm / P::B;
        const float2 w = m.src_ab[j];
        const float val = sm.amp * (w.x * sm.sin_t - w.y * sm.cos_t);
#pragma unroll
        for (int r = 0; r < P::R2; ++r) {
#pragma unroll
          for (int k1 = 0; k1 < P::A; ++k1) {
            if (k2 == t + P::T * r && k1 == k1s) {
              if (zz & 1) X[r][k1].y += val;
              else X[r][k1].x += val;
            }
          }
        }
      }
    }
    // store p, convergence peak, record
    unsigned pk = 0u;
    const bool in_conv_line = valid && x >= m.conv[0] && x < m.conv[1] && y >= m.conv[2] && y < m.conv[3];
    const bool in_rec_line = valid && sm.record && x >= m.rec[0] && x < m.rec[1] && y >= m.rec[2] &&
                             y < m.rec[3];
    const int ry = m.rec[3] - m.rec[2], rz = m.rec[5] - m.rec[4];
    const long rline = ((long)(x - m.rec[0]) * ry + (y - m.rec[2])) * rz - m.rec[4];
#pragma unroll
    for (int r = 0; r < P::R2; ++r) {
      const int k2 = t + P::T * r;
      if (k2 < P::B) {
#pragma unroll
        for (int k1 = 0; k1 < P::A; ++k1) {
          const int mm = k2 + P::B * k1;
          if (valid) {
            const float2 pv = X[r][k1];
            pl[mm] = pv;
#pragma unroll
            for (int e = 0; e < 2; ++e) {
              const int zz = 2 * mm + e;
              const float v = e ? pv.y : pv.x;
              if (in_conv_line && zz >= m.conv[4] && zz < m.conv[5]) pk = max(pk, __float_as_uint(fabsf(v)));
              if (in_rec_line && zz >= m.rec[4] && zz < m.rec[5]) {
                const long ri = rline + zz;
                for (int h = 0; h < m.nh; ++h) {
                  const float2 eh = sm.eh[h];
                  float2 s = m.acc[h][ri], cp = m.comp[h][ri];
                  const float yx = v * eh.x - cp.x, yy = v * eh.y - cp.y;
                  const float tx = s.x + yx, ty = s.y + yy;
                  cp.x = (tx - s.x) - yx;
                  cp.y = (ty - s.y) - yy;
                  m.acc[h][ri] = make_float2(tx, ty);
                  m.comp[h][ri] = cp;
                }
                m.pmax[ri] = fmaxf(m.pmax[ri], fabsf(v));
              }
            }
          }
        }
      }
    }
    if (sm.peak_idx >= 0) {
      for (int off = 32; off > 0; off >>= 1) pk = max(pk, (unsigned)__shfl_xor((int)pk, off));
      if ((threadIdx.x & 63) == 0 && pk) atomicMax(m.peaks + sm.peak_idx, pk);
    }
  } else {
#pragma unroll
    for (int r = 0; r < P::R2; ++r) {
      const int k2 = t + P::T * r;
#pragma unroll
      for (int k1 = 0; k1 < P::A; ++k1) {
        const int mm = (k2 < P::B) ? k2 + P::B * k1 : 0;
        X[r][k1] = valid ? pl[mm] : make_float2(0.f, 0.f);
      }
    }
  }
  z_r2c_store<M>(t, X, lb, lbl, tw, rt, m.sp + lsafe * (M + 1), valid);
}

}  // namespace cst
