// Pass kernels of the opt-in staggered scheme (session scheme CST_SCHEME_STAGGERED).
//
// First-order fluid equations on a staggered grid with k-space corrected, half-cell
// shifted spectral derivatives (Tabei, Mast and Waag, JASA 111(1), 2002), fractional-
// Laplacian power-law absorption and dispersion (Treeby and Cox, JASA 127(5), 2010) and a
// split-field PML (Berenger). Per step n, for the axes a = x, y, z:
//   u_a   <- s'_a (s'_a u_a - dt / rho0'_a  d+_a p)           d+_a = IFFT(i k_a kappa e^{+i k_a dx/2} FFT)
//   rho_a <- s_a (s_a rho_a - dt (rho0 + 2 m rho) d-_a u_a)   d-_a = IFFT(i k_a kappa e^{-i k_a dx/2} FFT)
//   rho_a[src] += q(n) / (nd c^2)
//   p = c^2 (rho + tau L1 - eta L2 + m (beta - 1) rho^2 / rho0),  rho = sum_a rho_a
//   L1 = IFFT(|k|^(y-2) FFT(rho0 sum_a d-_a u_a)),  L2 = IFFT(|k|^(y-1) FFT(rho))
// The upstream solver has no such scheme (its step, engine.py:458-482, is collocated with
// exp(-alpha c dt) absorption and a multiplicative sponge; power-law absorption is its
// unstarted milestone M16, MILESTONES.md:1201).
//
// Launch sequence of one step, on the half-spectrum layout of the reference scheme
// (kspace_types.h); the velocity state is kept as its z-spectrum like there:
//   sxgrad   x-FFT of p-hat, A = f_x i kappa / N P, Q = i kappa / N P, inverse x-FFT -> sg0, sg1
//   sygrad   sg0 = IFFT_y(A), sg1 = IFFT_y(f_y Q), sg2 = f_z IFFT_y(Q)
//   szu      per axis: z-C2R of sg[a] and su[a], velocity update, z-R2C into su[a]
//   sydiv    per axis: sg[a] = conj(f_a) FFT_y(su[a]) (f_x is applied in sxdiv)
//   sxdiv    per axis: sg[a] = IFFT_x(i kappa / N [conj(f_x)] FFT_x(sg[a]))
//   y inv    (3 spectra)
//   szrho    z-C2R of the three divergence terms, density update, source; linear
//            absorption-free members finish p here (peak, record, z-R2C -> sp);
//            absorbing members write rho and z-R2C dt rho0 sum_a d_a and rho -> sg0, sg1
//   [y fwd (2), sxabs (|k|^(y-2) / N, |k|^(y-1) / N), y inv (2), szabs: p, peak, record,
//    z-R2C -> sp]                                                       absorbing members
//   y fwd    (1 spectrum)
// with f_a(k) = k_a exp(+i k_a dx / 2) (the half-cell shift; its conjugate shifts back).
#pragma once
#include "kspace_kernels.h"

namespace cst {

// float2 table [0, L) staged through registers into LDS, like TwStage
template <int L, int NT>
struct C2Stage {
  static constexpr int Q = (L + NT - 1) / NT;
  float2 v[Q];
  __device__ __forceinline__ void load(gptr<const float2> src) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i = threadIdx.x + q * NT;
      v[q] = src[i < L ? i : L - 1];
    }
  }
  __device__ __forceinline__ void put(float2* dst) const {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i = threadIdx.x + q * NT;
      if (i < L) dst[i] = v[q];
    }
  }
};

// store a transformed column X (bin k2 + B k1 held as X[r][k1]) times a complex scalar
template <int L, int T>
__device__ __forceinline__ void y_store_c(int t, bool valid, gptr<float2> dst, long base, long stride,
                                          const float2 (&X)[Split<L, T>::R2][Split<L, T>::A], float2 f) {
  using S = Split<L, T>;
  if (!valid) return;
#pragma unroll
  for (int r = 0; r < S::R2; ++r) {
    const int k2 = t + T * r;
    if (k2 < S::B) {
#pragma unroll
      for (int k1 = 0; k1 < S::A; ++k1) dst[base + (long)(k2 + S::B * k1) * stride] = cmul(X[r][k1], f);
    }
  }
}

// store a strided-pass column V (natural order m2 + A j held as V[r][j])
template <int L, int T>
__device__ __forceinline__ void x_store(int t, bool valid, gptr<float2> dst, long base, long stride,
                                        const float2 (&V)[Split<L, T>::R1][Split<L, T>::B]) {
  using S = Split<L, T>;
  if (!valid) return;
#pragma unroll
  for (int r = 0; r < S::R1; ++r) {
    const int m2 = t + T * r;
    if (m2 < S::A) {
#pragma unroll
      for (int j = 0; j < S::B; ++j) dst[base + (long)(m2 + S::A * j) * stride] = V[r][j];
    }
  }
}

// ================================================================ strided passes
// gradient: sg[0] = IFFT_x(f_x i kappa / N FFT_x(sp)), sg[1] = IFFT_x(i kappa / N FFT_x(sp)).
// Block 0 of each member advances the member's step clock (as k_xgrad does).
template <int L>
__global__ void __launch_bounds__((Plan<L>::TC * Plan<L>::T)) k_sxgrad(StepArgs a, const float2* __restrict__ twg) {
  using P = Plan<L>;
  constexpr int T = P::T;
  using S = Split<L, T>;
  constexpr int TC = P::TC;
  __shared__ float2 tile[L * TC];
  __shared__ float2 tw[L];
  __shared__ float2 kxt[L];
  __shared__ float2 fxt[L];
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      const gptr<long long> ck = G(m.clock);
      const long long n = ck[0];
      ck[0] = n + 1;
      G(m.stepv)[0] = step_scalars(m, n, ck[1]);
    }
    return;
  }
  TwStage<L, TC * T> tws;
  tws.load(twg);
  KxStage<L, TC * T> kxs;
  kxs.load(m);
  C2Stage<L, TC * T> fxs;
  fxs.load(G(m.stg.kfac));
  const int col = threadIdx.x % TC, t = threadIdx.x / TC;
  const Col cg = x_column<TC>(a.s, col, (int)blockIdx.x - 1);
  const long C = cg.stride;
  const long base = cg.base;
  float2 V[S::R1][S::B];
  const gptr<const float2> src = G(m.sp);
  in_load<L, T>(t, [&](int n) { return ld(src, base + n * C); }, V);
  tws.put(tw);
  kxs.put(kxt);
  fxs.put(fxt);
  __syncthreads();
  TileAt<TC> at{tile, col};
  in_compute<L, T, -1>(t, V, at, tw);
  __syncthreads();
  float2 X[S::R2][S::A];
  in_finish<L, T, -1>(t, at, X);
  const XCol c = x_col(m, a.s, cg.plane, cg.zb, kxt);
  float2 Z[S::R2][S::A];
#pragma unroll
  for (int r = 0; r < S::R2; ++r) {
    const int k2 = t + T * r;
#pragma unroll
    for (int k1 = 0; k1 < S::A; ++k1) {
      const int kxi = (k2 < S::B) ? k2 + S::B * k1 : 0;
      const float kp = c.kappa(kxi);
      X[r][k1] = make_float2(-kp * X[r][k1].y, kp * X[r][k1].x);  // i kappa / N X
      Z[r][k1] = cmul(fxt[kxi], X[r][k1]);
    }
  }
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    __syncthreads();
    out_compute<L, T, +1>(t, pass == 0 ? Z : X, at, tw);
    __syncthreads();
    out_finish<L, T, +1>(t, at, V);
    x_store<L, T>(t, cg.valid, G(m.sg[pass]), base, C, V);
  }
}

// sg0 <- IFFT_y(A), sg1 <- IFFT_y(f_y Q), sg2 <- f_z IFFT_y(Q) (f_z: the column's constant)
template <int L>
__global__ void __launch_bounds__((Plan<L>::TC * Plan<L>::T)) k_sygrad(StepArgs a, const float2* __restrict__ twg) {
  using P = Plan<L>;
  using S = Split<L, P::T>;
  constexpr int TC = P::TC;
  constexpr int T = P::T;
  __shared__ float2 tile[L * TC];
  __shared__ float2 tw[L];
  __shared__ float2 fyt[L];
  TwStage<L, TC * T> tws;
  tws.load(twg);
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  C2Stage<L, TC * T> fys;
  fys.load(G(m.stg.kfac) + a.s.nx);
  const int col = threadIdx.x % TC, t = threadIdx.x / TC;
  const Col cg = y_column<TC>(a.s, col);
  const long S_ = cg.stride;
  const long base = cg.base;
  const gptr<float2> g0 = G(m.sg[0]), g1 = G(m.sg[1]), g2 = G(m.sg[2]);
  float2 VA[S::R1][S::B], VQ[S::R1][S::B];
  in_load<L, T>(t, [&](int n) { return ld(g0, base + n * S_); }, VA);
  const float2 fz = G(m.stg.kfac)[a.s.nx + a.s.ny + cg.zb];
  tws.put(tw);
  fys.put(fyt);
  __syncthreads();
  in_load<L, T>(t, [&](int n) { return ld(g1, base + n * S_); }, VQ);
  TileAt<TC> at{tile, col};
  float2 X[S::R2][S::A];
  in_compute<L, T, +1>(t, VA, at, tw);
  __syncthreads();
  in_finish<L, T, +1>(t, at, X);
  y_store<L, T>(t, cg.valid, g0, base, S_, X, 1.f);
  // f_y Q: input bin n = n1 + A n2 of V[r][n2] (n1 = t + T r)
#pragma unroll
  for (int r = 0; r < S::R1; ++r) {
    const int n1 = cmin(t + T * r, S::A - 1);
#pragma unroll
    for (int n2 = 0; n2 < S::B; ++n2) VA[r][n2] = cmul(fyt[n1 + S::A * n2], VQ[r][n2]);
  }
  __syncthreads();
  in_compute<L, T, +1>(t, VA, at, tw);
  __syncthreads();
  in_finish<L, T, +1>(t, at, X);
  y_store<L, T>(t, cg.valid, g1, base, S_, X, 1.f);
  __syncthreads();
  in_compute<L, T, +1>(t, VQ, at, tw);
  __syncthreads();
  in_finish<L, T, +1>(t, at, X);
  y_store_c<L, T>(t, cg.valid, g2, base, S_, X, fz);
}

// one velocity component per blockIdx.y = a: sg[a] <- FFT_y(su[a]) times conj(f_y) (a = 1,
// along the transformed axis) or conj(f_z) (a = 2, the column's constant)
template <int L>
__global__ void __launch_bounds__((Plan<L>::TC * Plan<L>::T)) k_sydiv(StepArgs a, const float2* __restrict__ twg) {
  using P = Plan<L>;
  using S = Split<L, P::T>;
  constexpr int TC = P::TC;
  constexpr int T = P::T;
  __shared__ float2 tile[L * TC];
  __shared__ float2 tw[L];
  __shared__ float2 fyt[L];
  TwStage<L, TC * T> tws;
  tws.load(twg);
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  const int ax = blockIdx.y;
  C2Stage<L, TC * T> fys;
  fys.load(G(m.stg.kfac) + a.s.nx);
  const int col = threadIdx.x % TC, t = threadIdx.x / TC;
  const Col cg = y_column<TC>(a.s, col);
  const long S_ = cg.stride;
  const long base = cg.base;
  const gptr<const float2> src = G(m.su[ax]);
  float2 V[S::R1][S::B];
  in_load<L, T>(t, [&](int n) { return ld(src, base + n * S_); }, V);
  const float2 fz = G(m.stg.kfac)[a.s.nx + a.s.ny + cg.zb];
  tws.put(tw);
  fys.put(fyt);
  __syncthreads();
  TileAt<TC> at{tile, col};
  in_compute<L, T, -1>(t, V, at, tw);
  __syncthreads();
  float2 X[S::R2][S::A];
  in_finish<L, T, -1>(t, at, X);
  if (ax == 1) {
#pragma unroll
    for (int r = 0; r < S::R2; ++r) {
      const int k2 = cmin(t + T * r, S::B - 1);
#pragma unroll
      for (int k1 = 0; k1 < S::A; ++k1) X[r][k1] = cmulconj(X[r][k1], fyt[k2 + S::B * k1]);
    }
  }
  y_store_c<L, T>(t, cg.valid, G(m.sg[ax]), base, S_, X,
                  ax == 2 ? make_float2(fz.x, -fz.y) : make_float2(1.f, 0.f));
}

// one divergence term per blockIdx.y = a: sg[a] <- IFFT_x(i kappa / N [conj(f_x)] FFT_x(sg[a])),
// conj(f_x) for a = 0 only (the y and z factors were applied by k_sydiv)
template <int L>
__global__ void __launch_bounds__((Plan<L>::TC * Plan<L>::T)) k_sxdiv(StepArgs a, const float2* __restrict__ twg) {
  using P = Plan<L>;
  constexpr int T = P::T;
  using S = Split<L, T>;
  constexpr int TC = P::TC;
  __shared__ float2 tile[L * TC];
  __shared__ float2 tw[L];
  __shared__ float2 kxt[L];
  __shared__ float2 fxt[L];
  TwStage<L, TC * T> tws;
  tws.load(twg);
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  const int ax = blockIdx.y;
  KxStage<L, TC * T> kxs;
  kxs.load(m);
  C2Stage<L, TC * T> fxs;
  fxs.load(G(m.stg.kfac));
  const int col = threadIdx.x % TC, t = threadIdx.x / TC;
  const Col cg = x_column<TC>(a.s, col);
  const long C = cg.stride;
  const long base = cg.base;
  const gptr<float2> data = G(m.sg[ax]);
  float2 V[S::R1][S::B];
  in_load<L, T>(t, [&](int n) { return ld(data, base + n * C); }, V);
  tws.put(tw);
  kxs.put(kxt);
  fxs.put(fxt);
  __syncthreads();
  TileAt<TC> at{tile, col};
  in_compute<L, T, -1>(t, V, at, tw);
  __syncthreads();
  float2 X[S::R2][S::A];
  in_finish<L, T, -1>(t, at, X);
  const XCol c = x_col(m, a.s, cg.plane, cg.zb, kxt);
#pragma unroll
  for (int r = 0; r < S::R2; ++r) {
    const int k2 = t + T * r;
#pragma unroll
    for (int k1 = 0; k1 < S::A; ++k1) {
      const int kxi = (k2 < S::B) ? k2 + S::B * k1 : 0;
      const float kp = c.kappa(kxi);
      float2 v = make_float2(-kp * X[r][k1].y, kp * X[r][k1].x);
      if (ax == 0) v = cmulconj(v, fxt[kxi]);
      X[r][k1] = v;
    }
  }
  __syncthreads();
  out_compute<L, T, +1>(t, X, at, tw);
  __syncthreads();
  out_finish<L, T, +1>(t, at, V);
  x_store<L, T>(t, cg.valid, data, base, C, V);
}

// fractional Laplacians, one per blockIdx.y = b: sg[b] <- IFFT_x(|k|^(2 e_b) / N FFT_x(sg[b]))
// with |k|^2 = kx^2 + ky^2 + kz^2 (full wavenumbers), e_0 = (y - 2) / 2, e_1 = (y - 1) / 2,
// and the k = 0 bin zeroed
template <int L>
__global__ void __launch_bounds__((Plan<L>::TC * Plan<L>::T)) k_sxabs(StepArgs a, const float2* __restrict__ twg) {
  using P = Plan<L>;
  constexpr int T = P::T;
  using S = Split<L, T>;
  constexpr int TC = P::TC;
  __shared__ float2 tile[L * TC];
  __shared__ float2 tw[L];
  __shared__ float2 kxt[L];
  TwStage<L, TC * T> tws;
  tws.load(twg);
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  const int b = blockIdx.y;
  KxStage<L, TC * T> kxs;
  kxs.load(m);
  const int col = threadIdx.x % TC, t = threadIdx.x / TC;
  const Col cg = x_column<TC>(a.s, col);
  const long C = cg.stride;
  const long base = cg.base;
  const gptr<float2> data = G(m.sg[b]);
  float2 V[S::R1][S::B];
  in_load<L, T>(t, [&](int n) { return ld(data, base + n * C); }, V);
  const gptr<const float> kt = G(m.ktab);
  const float kyf = kt[2 * a.s.nx + cg.plane], kzf = kt[2 * a.s.nx + 2 * a.s.ny + cg.zb];
  const float kyz2 = kyf * kyf + kzf * kzf;
  const float e = b == 0 ? m.stg.e1 : m.stg.e2;
  const float inv_n = m.inv_n;
  tws.put(tw);
  kxs.put(kxt);
  __syncthreads();
  TileAt<TC> at{tile, col};
  in_compute<L, T, -1>(t, V, at, tw);
  __syncthreads();
  float2 X[S::R2][S::A];
  in_finish<L, T, -1>(t, at, X);
#pragma unroll
  for (int r = 0; r < S::R2; ++r) {
    const int k2 = t + T * r;
#pragma unroll
    for (int k1 = 0; k1 < S::A; ++k1) {
      const int kxi = (k2 < S::B) ? k2 + S::B * k1 : 0;
      const float kx = kxt[kxi].x;
      const float q = kx * kx + kyz2;
      const float f = q > 0.f ? exp2f(e * log2f(q)) * inv_n : 0.f;
      X[r][k1] = cscale(X[r][k1], f);
    }
  }
  __syncthreads();
  out_compute<L, T, +1>(t, X, at, tw);
  __syncthreads();
  out_finish<L, T, +1>(t, at, V);
  x_store<L, T>(t, cg.valid, data, base, C, V);
}

// ================================================================ z passes
// Pressure sink of the z passes that finish p: convergence peak over the member's
// convergence box and, on a record step, the Kahan phasor sums and p_max over its record
// box (the epilogue of k_zp). Descriptor fields are read once up front: a global store
// may alias the descriptor table as far as the compiler knows.
struct PSink {
  int cv[6], rb[6];
  int nh, ry, rz, record;
  float2 eh[kMaxHarmonics];
  gptr<float2> acc[kMaxHarmonics], comp[kMaxHarmonics];
  gptr<float> pmax;
  unsigned pk;
  __device__ __forceinline__ void init(const MemberDev& m, const StepMember& sm) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      cv[q] = m.conv[q];
      rb[q] = m.rec[q];
    }
    nh = m.nh;
#pragma unroll
    for (int h = 0; h < kMaxHarmonics; ++h) {
      acc[h] = G(m.acc[h]);
      comp[h] = G(m.comp[h]);
      eh[h] = sm.eh[h];
    }
    pmax = G(m.pmax);
    ry = rb[3] - rb[2];
    rz = rb[5] - rb[4];
    record = sm.record;
    pk = 0u;
  }
  // elements z = 2 mm, 2 mm + 1 of line (x, y)
  __device__ __forceinline__ void put(int x, int y, int mm, float2 pv) {
    if (x >= cv[0] && x < cv[1] && y >= cv[2] && y < cv[3]) {
      const int z0 = 2 * mm;
      if (z0 >= cv[4] && z0 < cv[5]) pk = max(pk, __float_as_uint(fabsf(pv.x)));
      if (z0 + 1 >= cv[4] && z0 + 1 < cv[5]) pk = max(pk, __float_as_uint(fabsf(pv.y)));
    }
    if (record && x >= rb[0] && x < rb[1] && y >= rb[2] && y < rb[3]) {
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int zz = 2 * mm + e;
        const float v = e ? pv.y : pv.x;
        if (zz >= rb[4] && zz < rb[5]) {
          const long ri = ((long)(x - rb[0]) * ry + (y - rb[2])) * rz + (zz - rb[4]);
#pragma unroll
          for (int h = 0; h < kMaxHarmonics; ++h) {
            if (h >= nh) break;
            const float2 sacc = acc[h][ri];
            float2 cp = comp[h][ri];
            const float yx = v * eh[h].x - cp.x, yy = v * eh[h].y - cp.y;
            const float tx = sacc.x + yx, ty = sacc.y + yy;
            cp.x = (tx - sacc.x) - yx;
            cp.y = (ty - sacc.y) - yy;
            acc[h][ri] = make_float2(tx, ty);
            comp[h][ri] = cp;
          }
          pmax[ri] = fmaxf(pmax[ri], fabsf(v));
        }
      }
    }
  }
  // one atomic per wave into the period's slots (spread by wave id)
  template <int THREADS>
  __device__ __forceinline__ void flush(unsigned* peaks, int peak_idx) {
    if (peak_idx < 0) return;
    for (int off = 32; off > 0; off >>= 1) pk = max(pk, (unsigned)__shfl_xor((int)pk, off));
    if ((threadIdx.x & 63) == 0 && pk) {
      const int slot = (blockIdx.x * ((THREADS + 63) / 64) + (threadIdx.x >> 6)) % kPeakSlots;
      atomicMax(peaks + (long)peak_idx * kPeakWords + slot * kPeakStride, pk);
    }
  }
};

// velocity update, one axis per block (block -> (line group, axis) as in k_zu):
// u_a = s'_a (s'_a u_a - dt / rho0'_a g_a), s'_a the PML factor at staggered points along
// axis a only (a table with ones along the other axes), g_a = d+_a p from sg[a]
template <int M>
__global__ void __launch_bounds__((ZPlan<M, 0>::THREADS))
    k_szu(StepArgs a, const float2* __restrict__ twg, const float2* __restrict__ rtg) {
  using Z = ZPlan<M, 0>;
  using CN = ZCount<M, 0>;
  __shared__ ZLds<M, 0> s;
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  const int nlines = a.s.nx * a.s.ny;
  const int ngroups = (nlines + Z::LPB - 1) / Z::LPB;
  const int b = blockIdx.x;
  const int g0 = (b / 24) * 8 + (b % 8);
  const int ax = (b / 8) % 3;
  if (g0 >= ngroups) return;
  const int g = ngroups - 1 - g0;
  const long line0 = (long)g * Z::LPB;
  const int nl = cmin(Z::LPB, nlines - (int)line0);
  const int ll = threadIdx.x / Z::ZT, t = threadIdx.x % Z::ZT;
  const bool lvalid = ll < nl;
  const long e0 = line0 * M;
  const gptr<float2> gspec = G(m.sg[ax]);
  const gptr<float2> uspec = G(m.su[ax]);
  const float inz = m.inv_nz;
  ZStage<M, 0> stg;
  z_stage_load<M, 0>(twg, rtg, m.stg.pml + (1 + ax) * m.stg.pml_stride, a.s, line0, nl, stg);
  float2 sr[CN::EPT], ur[CN::EPT], dto[CN::EPT], srn, urn;
  z_spec_load<M, 0>(gspec, line0, nl, a.s, sr, srn);
  z_spec_load<M, 0>(uspec, line0, nl, a.s, ur, urn);
  z_stage_put<M, 0>(stg, a.s, line0, nl, s);
  z_spec_put<M, 0>(nl, sr, srn, s);
  __syncthreads();
  z_c2r<M, 0>(ll, t, lvalid, s);
  z_vol_load<M, 0>(G(reinterpret_cast<const float2*>(m.stg.dtrho[ax])) + e0, nl, dto);
  float2 gg[CN::EPT];
#pragma unroll
  for (int j = 0; j < CN::EPT; ++j) {
    const int l0 = ZElem<M, 0>::line(j), mm = ZElem<M, 0>::pos(j, l0);
    const int l = l0 < nl ? l0 : nl - 1;
    gg[j] = s.a[l * Z::LSA + mm];
  }
  __syncthreads();
  z_spec_put<M, 0>(nl, ur, urn, s);
  __syncthreads();
  z_c2r<M, 0>(ll, t, lvalid, s);
#pragma unroll
  for (int j = 0; j < CN::EPT; ++j) {
    const int l = ZElem<M, 0>::line(j), mm = ZElem<M, 0>::pos(j, l);
    if (l < nl) {
      const float sxy = s.sxy[l];
      const float2 s2 = s.sz[mm];
      const float fx = sxy * s2.x, fy = sxy * s2.y;
      float2 uv = s.a[l * Z::LSA + mm];
      uv.x = fx * (fx * (uv.x * inz) - dto[j].x * gg[j].x);
      uv.y = fy * (fy * (uv.y * inz) - dto[j].y * gg[j].y);
      s.a[l * Z::LSA + mm] = uv;
    }
  }
  __syncthreads();
  z_r2c<M, 0>(ll, t, lvalid, s, uspec + s.row[lvalid ? ll : 0], uspec + a.s.so + line0 + (lvalid ? ll : 0));
}

// PML factors of the density update at grid points: s_x[x], s_y[y] per line, s_z pairs
template <int M>
struct RhoLds {
  float sx[ZPlan<M, 0>::LPB];
  float sy[ZPlan<M, 0>::LPB];
};

// density update: d_a = d-_a u_a from the three divergence spectra sg[a],
// rho_a = s_a (s_a rho_a - (dt rho0 + 2 m dt rho) d_a), the mass source into rho_a for the
// grid's axes, rho = sum_a rho_a. ABS = false: p = c^2 rho (+ c^2 m (beta - 1) rho^2 / rho0),
// peak, record and z-R2C of p into sp. ABS = true: rho goes to p's buffer for k_szabs and
// dt rho0 sum_a d_a and rho are z-R2C'd into sg0, sg1 for the fractional Laplacians.
// The components are streamed one at a time (load, update, source, store), so that only
// one of them, the running sum and the weight are held in registers; the source adds to
// each component right after its update, which is where the oracle's order puts it too
// (the update of one component does not read the others').
template <int M, bool ABS>
__global__ void __launch_bounds__((ZPlan<M, 0>::THREADS))
    k_szrho(StepArgs a, const float2* __restrict__ twg, const float2* __restrict__ rtg) {
  using Z = ZPlan<M, 0>;
  using CN = ZCount<M, 0>;
  __shared__ ZLds<M, 0> s;
  __shared__ RhoLds<M> rs;
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  const StepMember sm = member_step(m);
  const int nlines = a.s.nx * a.s.ny;
  const long line0 = (long)(gridDim.x - 1 - blockIdx.x) * Z::LPB;
  const int nl = cmin(Z::LPB, nlines - (int)line0);
  const int ll = threadIdx.x / Z::ZT, t = threadIdx.x % Z::ZT;
  const bool lvalid = ll < nl;
  const long e0 = line0 * M;
  const bool nonlin = m.stg.nl2dt != nullptr;
  const int src_axes = m.stg.src_axes;
  const gptr<float2> rhog[3] = {G(reinterpret_cast<float2*>(m.stg.rho[0])) + e0,
                                G(reinterpret_cast<float2*>(m.stg.rho[1])) + e0,
                                G(reinterpret_cast<float2*>(m.stg.rho[2])) + e0};
  const gptr<const float2> dto0g = G(reinterpret_cast<const float2*>(m.stg.dtrho0)) + e0;
  ZStage<M, 0> stg;
  z_stage_load<M, 0>(twg, rtg, m.stg.pml, a.s, line0, nl, stg);
  float2 sr[CN::EPT], srn, fac[CN::EPT];
  z_spec_load<M, 0>(G(m.sg[0]), line0, nl, a.s, sr, srn);
  z_vol_load<M, 0>(dto0g, nl, fac);
  const gptr<const int> sptr = G(m.src_ptr);
  const int sp_l = sptr[line0 + ((int)threadIdx.x <= nl ? (int)threadIdx.x : nl)];
  z_stage_put<M, 0>(stg, a.s, line0, nl, s);
  if ((int)threadIdx.x < nl) {
    const gptr<const float> tab = G(m.stg.pml) + a.s.nz;
    rs.sx[threadIdx.x] = tab[s.lx[threadIdx.x]];
    rs.sy[threadIdx.x] = tab[a.s.nx + s.ly[threadIdx.x]];
  }
  if (threadIdx.x <= nl) s.sptr[threadIdx.x] = sp_l;
  for (int i = (int)threadIdx.x + Z::THREADS; i <= nl; i += Z::THREADS) s.sptr[i] = sptr[line0 + i];
  z_spec_put<M, 0>(nl, sr, srn, s);
  __syncthreads();
  z_spec_load<M, 0>(G(m.sg[1]), line0, nl, a.s, sr, srn);
  if (nonlin) {
    // the weight takes the density before this step's update: sum the old components
    float2 os[CN::EPT], o[CN::EPT];
    z_vol_load<M, 0>(rhog[0], nl, os);
#pragma unroll
    for (int ax = 1; ax < 3; ++ax) {
      z_vol_load<M, 0>(rhog[ax], nl, o);
#pragma unroll
      for (int j = 0; j < CN::EPT; ++j) {
        os[j].x += o[j].x;
        os[j].y += o[j].y;
      }
    }
    z_vol_load<M, 0>(G(reinterpret_cast<const float2*>(m.stg.nl2dt)) + e0, nl, o);
#pragma unroll
    for (int j = 0; j < CN::EPT; ++j) {
      fac[j].x += o[j].x * os[j].x;
      fac[j].y += o[j].y * os[j].y;
    }
  }
  const gptr<const int> srcz = G(m.src_z);
  const gptr<const float2> srcab = G(m.src_ab);
  float2 rsum[CN::EPT];               // rho = (rho_0 + rho_1) + rho_2
  float2 ds[ABS ? CN::EPT : 1];       // sum_a d_a (absorbing members)
  // axis a: C2R of its divergence term (in s.a), rho_a = s_a (s_a rho_a - fac d_a), the
  // source, the store of rho_a
  auto update = [&](int ax) {
    float2 r[CN::EPT];
    z_vol_load<M, 0>(rhog[ax], nl, r);
    z_c2r<M, 0>(ll, t, lvalid, s);
#pragma unroll
    for (int j = 0; j < CN::EPT; ++j) {
      const int l0 = ZElem<M, 0>::line(j), mm = ZElem<M, 0>::pos(j, l0);
      const int l = l0 < nl ? l0 : nl - 1;
      const float2 d = s.a[l * Z::LSA + mm];
      float fx, fy;
      if (ax == 0) {
        fx = fy = rs.sx[l];
      } else if (ax == 1) {
        fx = fy = rs.sy[l];
      } else {
        fx = s.sz[mm].x;
        fy = s.sz[mm].y;
      }
      float2 rv = make_float2(fx * (fx * r[j].x - fac[j].x * d.x), fy * (fy * r[j].y - fac[j].y * d.y));
      if ((src_axes >> ax) & 1) {
        // duplicates are refused at admission: at most one entry per element
        for (int e = s.sptr[l]; e < s.sptr[l + 1]; ++e) {
          const int zz = srcz[e];
          if ((zz >> 1) == mm) {
            const float2 w = srcab[e];
            const float v = sm.amp * (w.x * sm.sin_t - w.y * sm.cos_t);
            if (zz & 1) rv.y += v;
            else rv.x += v;
          }
        }
      }
      if (l0 < nl) rhog[ax][l * M + mm] = rv;
      if (ax == 0) {
        rsum[j] = rv;
      } else {
        rsum[j].x += rv.x;
        rsum[j].y += rv.y;
      }
      if constexpr (ABS) {
        if (ax == 0) ds[j] = d;
        else {
          ds[j].x += d.x;
          ds[j].y += d.y;
        }
      }
    }
    __syncthreads();  // s.a consumed
  };
  update(0);
  z_spec_put<M, 0>(nl, sr, srn, s);
  __syncthreads();
  z_spec_load<M, 0>(G(m.sg[2]), line0, nl, a.s, sr, srn);
  update(1);
  z_spec_put<M, 0>(nl, sr, srn, s);
  __syncthreads();
  const gptr<float2> pg = G(reinterpret_cast<float2*>(m.p)) + e0;
  if constexpr (!ABS) {
    float2 c2[CN::EPT], bh[CN::EPT];
    z_vol_load<M, 0>(G(reinterpret_cast<const float2*>(m.stg.c2)) + e0, nl, c2);
    if (nonlin) z_vol_load<M, 0>(G(reinterpret_cast<const float2*>(m.stg.c2bh)) + e0, nl, bh);
    update(2);
    PSink sink;
    sink.init(m, sm);
#pragma unroll
    for (int j = 0; j < CN::EPT; ++j) {
      const int l = ZElem<M, 0>::line(j), mm = ZElem<M, 0>::pos(j, l);
      if (l < nl) {
        const float2 rho = rsum[j];
        float2 pv = make_float2(c2[j].x * rho.x, c2[j].y * rho.y);
        if (nonlin) {
          pv.x += bh[j].x * rho.x * rho.x;
          pv.y += bh[j].y * rho.y * rho.y;
        }
        pg[l * M + mm] = pv;
        sink.put(s.lx[l], s.ly[l], mm, pv);
        s.a[l * Z::LSA + mm] = pv;
      }
    }
    sink.flush<Z::THREADS>(m.peaks, sm.peak_idx);
    __syncthreads();
    z_r2c<M, 0>(ll, t, lvalid, s, G(m.sp) + s.row[lvalid ? ll : 0], G(m.sp) + a.s.so + line0 + (lvalid ? ll : 0));
  } else {
    update(2);
    // dt rho0, the weight of sum_a d_a: the weight itself on linear members, reread on
    // nonlinear ones (whose weight carries the density term)
    if (nonlin) z_vol_load<M, 0>(dto0g, nl, fac);
#pragma unroll
    for (int j = 0; j < CN::EPT; ++j) {
      const int l = ZElem<M, 0>::line(j), mm = ZElem<M, 0>::pos(j, l);
      if (l < nl) {
        pg[l * M + mm] = rsum[j];
        s.a[l * Z::LSA + mm] = make_float2(fac[j].x * ds[j].x, fac[j].y * ds[j].y);
      }
    }
    __syncthreads();
    z_r2c<M, 0>(ll, t, lvalid, s, G(m.sg[0]) + s.row[lvalid ? ll : 0],
                G(m.sg[0]) + a.s.so + line0 + (lvalid ? ll : 0));
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CN::EPT; ++j) {
      const int l = ZElem<M, 0>::line(j), mm = ZElem<M, 0>::pos(j, l);
      if (l < nl) s.a[l * Z::LSA + mm] = rsum[j];
    }
    __syncthreads();
    z_r2c<M, 0>(ll, t, lvalid, s, G(m.sg[1]) + s.row[lvalid ? ll : 0],
                G(m.sg[1]) + a.s.so + line0 + (lvalid ? ll : 0));
  }
}

// pressure of absorbing members: L1 = C2R(sg0), L2 = C2R(sg1) (the fractional Laplacians
// of dt rho0 sum_a d_a and of rho), rho from p's buffer;
// p = c^2 rho + (c^2 tau / dt) L1 - c^2 eta L2 (+ c^2 m (beta - 1) rho^2 / rho0); peak,
// record, z-R2C of p into sp
template <int M>
__global__ void __launch_bounds__((ZPlan<M, 0>::THREADS))
    k_szabs(StepArgs a, const float2* __restrict__ twg, const float2* __restrict__ rtg) {
  using Z = ZPlan<M, 0>;
  using CN = ZCount<M, 0>;
  __shared__ ZLds<M, 0> s;
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  const StepMember sm = member_step(m);
  const int nlines = a.s.nx * a.s.ny;
  const long line0 = (long)(gridDim.x - 1 - blockIdx.x) * Z::LPB;
  const int nl = cmin(Z::LPB, nlines - (int)line0);
  const int ll = threadIdx.x / Z::ZT, t = threadIdx.x % Z::ZT;
  const bool lvalid = ll < nl;
  const long e0 = line0 * M;
  const bool nonlin = m.stg.nl2dt != nullptr;
  ZStage<M, 0> stg;
  z_stage_load<M, 0>(twg, rtg, m.stg.pml, a.s, line0, nl, stg);
  float2 sr[CN::EPT], srn;
  z_spec_load<M, 0>(G(m.sg[0]), line0, nl, a.s, sr, srn);
  const gptr<float2> pg = G(reinterpret_cast<float2*>(m.p)) + e0;
  float2 rho[CN::EPT], c2[CN::EPT], ct[CN::EPT], ce[CN::EPT], bh[CN::EPT], l1[CN::EPT];
  z_vol_load<M, 0>(pg, nl, rho);
  z_vol_load<M, 0>(G(reinterpret_cast<const float2*>(m.stg.c2)) + e0, nl, c2);
  z_stage_put<M, 0>(stg, a.s, line0, nl, s);
  z_spec_put<M, 0>(nl, sr, srn, s);
  __syncthreads();
  z_spec_load<M, 0>(G(m.sg[1]), line0, nl, a.s, sr, srn);
  z_c2r<M, 0>(ll, t, lvalid, s);
  z_vol_load<M, 0>(G(reinterpret_cast<const float2*>(m.stg.c2tau)) + e0, nl, ct);
  z_vol_load<M, 0>(G(reinterpret_cast<const float2*>(m.stg.c2eta)) + e0, nl, ce);
  if (nonlin) z_vol_load<M, 0>(G(reinterpret_cast<const float2*>(m.stg.c2bh)) + e0, nl, bh);
#pragma unroll
  for (int j = 0; j < CN::EPT; ++j) {
    const int l0 = ZElem<M, 0>::line(j), mm = ZElem<M, 0>::pos(j, l0);
    const int l = l0 < nl ? l0 : nl - 1;
    l1[j] = s.a[l * Z::LSA + mm];
  }
  __syncthreads();
  z_spec_put<M, 0>(nl, sr, srn, s);
  __syncthreads();
  z_c2r<M, 0>(ll, t, lvalid, s);
  PSink sink;
  sink.init(m, sm);
#pragma unroll
  for (int j = 0; j < CN::EPT; ++j) {
    const int l = ZElem<M, 0>::line(j), mm = ZElem<M, 0>::pos(j, l);
    if (l < nl) {
      const int i = l * M + mm;
      const float2 l2 = s.a[l * Z::LSA + mm];
      float2 pv;
      pv.x = (c2[j].x * rho[j].x + ct[j].x * l1[j].x) - ce[j].x * l2.x;
      pv.y = (c2[j].y * rho[j].y + ct[j].y * l1[j].y) - ce[j].y * l2.y;
      if (nonlin) {
        pv.x += bh[j].x * rho[j].x * rho[j].x;
        pv.y += bh[j].y * rho[j].y * rho[j].y;
      }
      pg[i] = pv;
      sink.put(s.lx[l], s.ly[l], mm, pv);
      s.a[l * Z::LSA + mm] = pv;
    }
  }
  sink.flush<Z::THREADS>(m.peaks, sm.peak_idx);
  __syncthreads();
  z_r2c<M, 0>(ll, t, lvalid, s, G(m.sp) + s.row[lvalid ? ll : 0], G(m.sp) + a.s.so + line0 + (lvalid ? ll : 0));
}

}  // namespace cst
