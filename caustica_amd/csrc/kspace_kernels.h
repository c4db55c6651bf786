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
// Every line FFT is a two-stage split L = A * B across T threads of a line, with the
// exchange and the twiddle table in LDS. Spectrum rows are padded to a multiple of 16
// bins (128 bytes) so that every column tile of the strided passes starts on a cache line.
#pragma once
#include <type_traits>
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
constexpr int cmax(int a, int b) { return a > b ? a : b; }

// L = A * B; stage 1: A DFTs of size B; stage 2: B DFTs of size A.
// Strided passes: T = max(A, B) threads per line, so every thread holds one DFT per
// stage (registers: 2 max(A, B) floats per buffer); TC columns per block, at least 16
// (a 128-byte row segment) and the tile L x TC in LDS at most 64 KiB. Tile variant TV
// (the autotuner's choice per shape, lengths >= 64): 0 that width, 1 half of it, 2 twice
// it (within the LDS bound).
constexpr int tile_cols(int L, int T, int TV) {
  const int tc = cmax(1, pow2_floor(cmin(cmax(256 / T, 16), 8192 / L)));
  if (L < 64 || TV == 0) return tc;
  if (TV == 1) return cmax(8, tc / 2);
  return cmin(tc * 2, pow2_floor(8192 / L));
}
template <int L, int TV = 0>
struct Plan {
  static constexpr int A = plan_a(L);
  static constexpr int B = L / A;
  static constexpr int T = cmax(A, B);
  static constexpr int TC = tile_cols(L, T, TV);
};

// Distribution of one line FFT over T threads, stage-1 count AA (the radix plan: by
// default the smallest divisor of L at least sqrt(L); the z passes may take its cofactor).
template <int L, int T, int AA = Plan<L>::A>
struct Split {
  static constexpr int A = AA;
  static constexpr int B = L / A;
  static_assert(A * B == L, "radix plan must factor the length");
  static constexpr int R1 = (A + T - 1) / T;  // stage-1 DFTs per thread
  static constexpr int R2 = (B + T - 1) / T;  // stage-2 DFTs per thread
};

// z passes: complex length M = nz / 2; ZT threads per line, LPB lines per block.
// Variant V: layout V & 3 and radix plan V >> 2. Layout 0: ZT = max(A, B) (every lane
// busy in the larger stage); 1: twice the threads per line (half the registers per
// thread); 2 (velocity pass only): the layout of 0 with half the lines per block. Radix
// plan 0: A = the smallest divisor of M at least sqrt(M); 1: its cofactor (the two
// stages swap lengths). The session picks one variant per shape.
constexpr int zlayout(int V) { return V & 3; }
constexpr int zradix_a(int M, int V) { return (V >> 2) ? M / plan_a(M) : plan_a(M); }
template <int M, int V>
struct ZPlan {
  static constexpr int A = zradix_a(M, V);
  static constexpr int B = M / A;
  static constexpr int W = zlayout(V) == 1 ? 2 : 1;  // threads per line, in units of max(A, B)
  static constexpr int ZT = cmin(64, W * cmax(A, B));
  static constexpr int LSA = M + 2;                 // natural-order line buffer (M + 1 bins)
  static constexpr int LSE = M + (M + A - 1) / A;   // exchange buffer, padded every A rows
  static constexpr int LPB0 = cmax(1, pow2_floor(cmin(W * 256 / ZT, 4480 / (LSA + LSE))));
  static constexpr int LPB = zlayout(V) == 2 ? cmax(1, LPB0 / 2) : LPB0;
  static constexpr int THREADS = LPB * ZT;
  // V = 2 with one-wave blocks: the natural and exchange line buffers share one LDS
  // array (two more barriers per transform, free for a single wave), which lets more
  // blocks share a CU's LDS where LDS rather than registers bounds the waves per SIMD
  static constexpr bool ALIAS = zlayout(V) == 2 && THREADS <= 64;
};

template <int DIR>
__device__ __forceinline__ float2 twmul(float2 v, float2 w) {
  return DIR > 0 ? cmul(v, w) : cmulconj(v, w);
}

// ---------------------------------------------------------------- stages
// stage "in": natural-order input x[n1 + A n2]  ->  X[k2 + B k1] held as X[r][k1], k2 = t + T r
template <int L, int T, int AA = Plan<L>::A, class Load>
__device__ __forceinline__ void in_load(int t, Load load,
                                        float2 (&V)[Split<L, T, AA>::R1][Split<L, T, AA>::B]) {
  using S = Split<L, T, AA>;
#pragma unroll
  for (int r = 0; r < S::R1; ++r) {
    const int n1 = t + T * r;
    if constexpr (T * S::R1 == S::A) {
      // every thread owns a sub-line: unconditional loads keep the wait counts exact
#pragma unroll
      for (int n2 = 0; n2 < S::B; ++n2) V[r][n2] = load(n1 + S::A * n2);
    } else {
#pragma unroll
      for (int n2 = 0; n2 < S::B; ++n2)
        V[r][n2] = (n1 < S::A) ? load(n1 + S::A * n2) : make_float2(0.f, 0.f);
    }
  }
}

// TT: the twiddle table is stored transposed, tw[k2 A + n1] = w^(n1 k2), for passes
// whose consecutive lanes hold consecutive n1 of one line (the z passes): the natural
// layout would have them read at stride k2 and collide on LDS banks.
template <int L, int T, int DIR, bool TT = false, int AA = Plan<L>::A, class At>
__device__ __forceinline__ void in_compute(int t, float2 (&V)[Split<L, T, AA>::R1][Split<L, T, AA>::B],
                                           At at, const float2* tw) {
  using S = Split<L, T, AA>;
#pragma unroll
  for (int r = 0; r < S::R1; ++r) {
    const int n1 = t + T * r;
    if (n1 < S::A) {
      Dft<S::B, DIR>::run(V[r]);
#pragma unroll
      for (int k2 = 0; k2 < S::B; ++k2)
        *at(n1 + S::A * k2) = twmul<DIR>(V[r][k2], tw[TT ? k2 * S::A + n1 : n1 * k2]);
    }
  }
}

template <int L, int T, int DIR, int AA = Plan<L>::A, class At>
__device__ __forceinline__ void in_finish(int t, At at,
                                          float2 (&X)[Split<L, T, AA>::R2][Split<L, T, AA>::A]) {
  using S = Split<L, T, AA>;
#pragma unroll
  for (int r = 0; r < S::R2; ++r) {
    const int k2 = t + T * r;
    if (k2 < S::B) {
#pragma unroll
      for (int n1 = 0; n1 < S::A; ++n1) X[r][n1] = *at(n1 + S::A * k2);
      Dft<S::A, DIR>::run(X[r]);
    }
  }
}

// Stage "in" finish shared by two threads per A-point DFT, for lines whose T threads
// number twice the B stage-2 DFTs (T = 2 B, A even; e.g. M = 128 = 16 x 8 on 16 threads,
// where the plain finish leaves half the lanes idle). Threads k2 and k2 + B take column
// k2 as one radix-2 decimation-in-frequency step: x[n] + x[n + A/2] (even outputs) and
// (x[n] - x[n + A/2]) W_A^n (odd outputs), then an (A/2)-point DFT. Thread t = k2 + h B
// ends with outputs k1 = 2 m + h, m < A/2, i.e. bins k2 + B k1 = t + T m.
template <int L, int T, int AA>
struct PairFinish {
  using S = Split<L, T, AA>;
  static constexpr bool OK = 2 * S::B == T && S::A % 2 == 0 && S::A >= 4;
  static constexpr int H = S::A / 2;
};

template <int DIR, int A, int N, class At>
__device__ __forceinline__ float2 pair_in(At at, int k2, float sg, bool odd) {
  const float2 x0 = *at(N + A * k2), x1 = *at(N + A / 2 + A * k2);
  // x0 + sg x1 with sg = +-1 rounds exactly as the sum or the difference
  const float2 v = make_float2(fmaf(sg, x1.x, x0.x), fmaf(sg, x1.y, x0.y));
  return odd ? twiddle_const<DIR, N, A>(v) : v;
}

template <int L, int T, int DIR, int AA, class At, int... N>
__device__ __forceinline__ void in_finish_pair_(int t, At at, float2 (&Y)[PairFinish<L, T, AA>::H],
                                                std::integer_sequence<int, N...>) {
  using S = Split<L, T, AA>;
  const bool odd = t >= S::B;
  const int k2 = odd ? t - S::B : t;
  const float sg = odd ? -1.f : 1.f;
  ((Y[N] = pair_in<DIR, S::A, N>(at, k2, sg, odd)), ...);
  Dft<S::A / 2, DIR>::run(Y);
}

template <int L, int T, int DIR, int AA, class At>
__device__ __forceinline__ void in_finish_pair(int t, At at, float2 (&Y)[PairFinish<L, T, AA>::H]) {
  static_assert(PairFinish<L, T, AA>::OK, "pair finish needs T = 2 B and A even");
  in_finish_pair_<L, T, DIR, AA>(t, at, Y, std::make_integer_sequence<int, PairFinish<L, T, AA>::H>{});
}

// stage "out": input X[k2 + B k1] as X[r][k1]  ->  natural x[m2 + A j] as V[r][j], m2 = t + T r
template <int L, int T, int DIR, class At>
__device__ __forceinline__ void out_compute(int t, float2 (&X)[Split<L, T>::R2][Split<L, T>::A], At at,
                                            const float2* tw) {
  using S = Split<L, T>;
#pragma unroll
  for (int r = 0; r < S::R2; ++r) {
    const int k2 = t + T * r;
    if (k2 < S::B) {
      float2 v[S::A];
#pragma unroll
      for (int i = 0; i < S::A; ++i) v[i] = X[r][i];
      Dft<S::A, DIR>::run(v);
#pragma unroll
      for (int m2 = 0; m2 < S::A; ++m2) *at(k2 + S::B * m2) = twmul<DIR>(v[m2], tw[k2 * m2]);
    }
  }
}

template <int L, int T, int DIR, class At>
__device__ __forceinline__ void out_finish(int t, At at, float2 (&V)[Split<L, T>::R1][Split<L, T>::B]) {
  using S = Split<L, T>;
#pragma unroll
  for (int r = 0; r < S::R1; ++r) {
    const int m2 = t + T * r;
    if (m2 < S::A) {
#pragma unroll
      for (int k2 = 0; k2 < S::B; ++k2) V[r][k2] = *at(k2 + S::B * m2);
      Dft<S::B, DIR>::run(V[r]);
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

// Line exchange buffer with one padding slot after every PAD rows: stage 2 of the split
// reads rows n1 + A k2 at stride A across threads, which would otherwise hit the same
// LDS banks.
template <int PAD>
struct LinePadAt {
  float2* base;
  __device__ __forceinline__ float2* operator()(int row) const { return base + row + row / PAD; }
};

// Twiddle table of a strided pass, staged through registers: written to LDS after the
// column loads are issued, so those are in flight while the block waits for the table.
template <int L, int NT>
struct TwStage {
  static constexpr int Q = (L + NT - 1) / NT;
  float2 v[Q];
  __device__ __forceinline__ void load(const float2* __restrict__ twg) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i = threadIdx.x + q * NT;
      v[q] = twg[i < L ? i : L - 1];
    }
  }
  __device__ __forceinline__ void put(float2* tw) const {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i = threadIdx.x + q * NT;
      if (i < L) tw[i] = v[q];
    }
  }
};

// sin(x) / x as its Taylor series in q = x^2 (alternating terms (-q)^n / (2n+1)!, up to
// n = 7): for the k-space correction x = (c_ref dt / 2) |k| <= (pi / 2) CFL sqrt(3) <= 1.64
// (CFL <= 0.6), so q <= 2.7 and the first omitted term is below 1e-12.
__device__ __forceinline__ float sinc_q(float q) {
  float r = -1.f / 1307674368000.f;  // -1/15!
  r = fmaf(r, q, 1.f / 6227020800.f);
  r = fmaf(r, q, -1.f / 39916800.f);
  r = fmaf(r, q, 1.f / 362880.f);
  r = fmaf(r, q, -1.f / 5040.f);
  r = fmaf(r, q, 1.f / 120.f);
  r = fmaf(r, q, -1.f / 6.f);
  return fmaf(r, q, 1.f);
}

// ================================================================ y pass
// In-place C2C along y of one or three half spectra. Block = (x, column tile of TC bins).
// Column geometry of a strided-pass block (the spectrum layout is described at Shape).
// Main blocks take TC consecutive bins k < nz/2 of one plane (y pass: an x plane,
// stride ms along y; x pass: a y row, stride ps along x); the last
// ceil(planes / TC) blocks take the Nyquist bin k = nz/2 of TC consecutive planes from
// the side plane. Invalid columns address a valid element and are never stored.
struct Col {
  long base, stride;
  int plane, zb;
  bool valid;
};
__host__ __device__ constexpr int strided_blocks(int planes, int nzh, int tc) {
  return planes * ((nzh - 1 + tc - 1) / tc) + (planes + tc - 1) / tc;
}
// y passes: the side blocks take the first block indices. Their columns are contiguous
// runs of the side plane and finish quickly; dispatched last they formed a partial final
// round of the grid (at 256^3: 2048 main + 16 side blocks over 1024 block slots), and
// the y passes took 5-7 % longer.
template <int TC>
__device__ __forceinline__ Col y_column(const Shape& s, int col) {
  const int M = s.nzh - 1, nt = (M + TC - 1) / TC, ns = (s.nx + TC - 1) / TC;
  const int b = (int)blockIdx.x < ns ? (int)blockIdx.x + s.nx * nt : (int)blockIdx.x - ns;
  Col c;
  if (b < s.nx * nt) {
    c.plane = b / nt;
    const int zb0 = (b - c.plane * nt) * TC + col;
    c.valid = zb0 < M;
    c.zb = c.valid ? zb0 : 0;
    c.base = (long)c.plane * s.ps + c.zb;
    c.stride = s.ms;
  } else {
    const int x0 = (b - s.nx * nt) * TC + col;
    c.valid = x0 < s.nx;
    c.plane = c.valid ? x0 : 0;
    c.zb = M;
    c.base = s.so + (long)c.plane * s.ny;
    c.stride = 1;
  }
  return c;
}
template <int TC>
__device__ __forceinline__ Col x_column(const Shape& s, int col, int b = blockIdx.x) {
  const int M = s.nzh - 1, nt = (M + TC - 1) / TC;
  Col c;
  if (b < s.ny * nt) {
    c.plane = b / nt;
    const int zb0 = (b - c.plane * nt) * TC + col;
    c.valid = zb0 < M;
    c.zb = c.valid ? zb0 : 0;
    c.base = (long)c.plane * s.ms + c.zb;
    c.stride = s.ps;
  } else {
    const int y0 = (b - s.ny * nt) * TC + col;
    c.valid = y0 < s.ny;
    c.plane = c.valid ? y0 : 0;
    c.zb = M;
    c.base = s.so + c.plane;
    c.stride = s.ny;
  }
  return c;
}

template <int L, int DIR, int TV = 0>
__global__ void __launch_bounds__((Plan<L, TV>::TC * Plan<L, TV>::T))
    k_ypass(StepArgs a, const float2* __restrict__ twg, int which) {
  using P = Plan<L, TV>;
  using S = Split<L, P::T>;
  constexpr int TC = P::TC;
  constexpr int T = P::T;
  __shared__ float2 tile[L * TC];
  __shared__ float2 tw[L];
  TwStage<L, TC * T> tws;
  tws.load(twg);
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  const gptr<float2> data = G(which == 0 ? m.sp : m.sg[blockIdx.y]);
  const int col = threadIdx.x % TC, t = threadIdx.x / TC;
  const Col cg = y_column<TC>(a.s, col);
  const bool valid = cg.valid;
  const long S_ = cg.stride;
  const long base = cg.base;
  float2 V[S::R1][S::B];
  // invalid columns load bin 0 of their row (a valid address) and are never stored:
  // unconditional loads let the wait-count pass wait for exactly what is consumed
  in_load<L, T>(t, [&](int n) { return ld(data, base + n * S_); }, V);
  tws.put(tw);
  __syncthreads();
  TileAt<TC> at{tile, col};
  in_compute<L, T, DIR>(t, V, at, tw);
  __syncthreads();
  float2 X[S::R2][S::A];
  in_finish<L, T, DIR>(t, at, X);
  if (valid) {
#pragma unroll
    for (int r = 0; r < S::R2; ++r) {
      const int k2 = t + T * r;
      if (k2 < S::B) {
#pragma unroll
        for (int k1 = 0; k1 < S::A; ++k1) data[base + (long)(k2 + S::B * k1) * S_] = X[r][k1];
      }
    }
  }
}

// The two 3-spectrum y passes carry the y and z derivative factors, which are constant
// along x, so the x passes move two spectra instead of three:
//   after xgrad (sg0 = A, the x-derivative spectrum; sg1 = Q = i kappa / N FFT_x(p-hat)):
//     k_ygrad: sg0 <- IFFT_y(A), sg1 <- IFFT_y(ky_d Q), sg2 <- kz_d IFFT_y(Q)
//   after zu (su0..su2 = U_x, U_y, U_z, the velocity state, real along y):
//     k_ydiv:  sg0 <- FFT_y(U_x), sg1 <- ky_d FFT_y(U_y) + kz_d FFT_y(U_z)
// (ky_d varies along the transformed axis and multiplies in the spectral domain; kz_d is
// the column's constant.)
template <int L, int NT>
struct KyStage {
  static constexpr int Q = (L + NT - 1) / NT;
  float v[Q];
  __device__ __forceinline__ void load(const MemberDev& m, const Shape& sh) {
    const gptr<const float> kyd = G(m.ktab) + 2 * sh.nx + sh.ny;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i = threadIdx.x + q * NT;
      v[q] = kyd[i < L ? i : L - 1];
    }
  }
  __device__ __forceinline__ void put(float* kyt) const {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i = threadIdx.x + q * NT;
      if (i < L) kyt[i] = v[q];
    }
  }
};

// store a transformed column X (bin k2 + B k1 held as X[r][k1]) times a scalar
template <int L, int T>
__device__ __forceinline__ void y_store(int t, bool valid, gptr<float2> dst, long base, long stride,
                                        const float2 (&X)[Split<L, T>::R2][Split<L, T>::A], float f) {
  using S = Split<L, T>;
  if (!valid) return;
#pragma unroll
  for (int r = 0; r < S::R2; ++r) {
    const int k2 = t + T * r;
    if (k2 < S::B) {
#pragma unroll
      for (int k1 = 0; k1 < S::A; ++k1)
        dst[base + (long)(k2 + S::B * k1) * stride] = make_float2(f * X[r][k1].x, f * X[r][k1].y);
    }
  }
}

template <int L, int TV = 0>
__global__ void __launch_bounds__((Plan<L, TV>::TC * Plan<L, TV>::T))
    k_ygrad(StepArgs a, const float2* __restrict__ twg) {
  using P = Plan<L, TV>;
  using S = Split<L, P::T>;
  constexpr int TC = P::TC;
  constexpr int T = P::T;
  __shared__ float2 tile[L * TC];
  __shared__ float2 tw[L];
  __shared__ float kyt[L];
  TwStage<L, TC * T> tws;
  tws.load(twg);
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  KyStage<L, TC * T> kys;
  kys.load(m, a.s);
  const int col = threadIdx.x % TC, t = threadIdx.x / TC;
  const Col cg = y_column<TC>(a.s, col);
  const bool valid = cg.valid;
  const int zb = cg.zb;
  const long S_ = cg.stride;
  const long base = cg.base;
  const gptr<float2> g0 = G(m.sg[0]), g1 = G(m.sg[1]), g2 = G(m.sg[2]);
  float2 VA[S::R1][S::B], VQ[S::R1][S::B];
  in_load<L, T>(t, [&](int n) { return ld(g0, base + n * S_); }, VA);  // see k_ypass
  const float kzd = G(m.ktab)[2 * a.s.nx + 2 * a.s.ny + a.s.nzh + zb];
  tws.put(tw);
  kys.put(kyt);
  __syncthreads();
  // issued after the barrier (the tables' wait covers every earlier load): the first
  // transform overlaps this load
  in_load<L, T>(t, [&](int n) { return ld(g1, base + n * S_); }, VQ);
  TileAt<TC> at{tile, col};
  float2 X[S::R2][S::A];
  in_compute<L, T, +1>(t, VA, at, tw);
  __syncthreads();
  in_finish<L, T, +1>(t, at, X);
  y_store<L, T>(t, valid, g0, base, S_, X, 1.f);
  // ky_d Q: input bin n = n1 + A n2 of V[r][n2] (n1 = t + T r)
#pragma unroll
  for (int r = 0; r < S::R1; ++r) {
    const int n1 = cmin(t + T * r, S::A - 1);
#pragma unroll
    for (int n2 = 0; n2 < S::B; ++n2) {
      const float f = kyt[n1 + S::A * n2];
      VA[r][n2] = make_float2(f * VQ[r][n2].x, f * VQ[r][n2].y);
    }
  }
  __syncthreads();
  in_compute<L, T, +1>(t, VA, at, tw);
  __syncthreads();
  in_finish<L, T, +1>(t, at, X);
  y_store<L, T>(t, valid, g1, base, S_, X, 1.f);
  __syncthreads();
  in_compute<L, T, +1>(t, VQ, at, tw);
  __syncthreads();
  in_finish<L, T, +1>(t, at, X);
  y_store<L, T>(t, valid, g2, base, S_, X, kzd);
}

template <int L, int TV = 0>
__global__ void __launch_bounds__((Plan<L, TV>::TC * Plan<L, TV>::T))
    k_ydiv(StepArgs a, const float2* __restrict__ twg) {
  using P = Plan<L, TV>;
  using S = Split<L, P::T>;
  constexpr int TC = P::TC;
  constexpr int T = P::T;
  __shared__ float2 tile[L * TC];
  __shared__ float2 tw[L];
  __shared__ float kyt[L];
  TwStage<L, TC * T> tws;
  tws.load(twg);
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  KyStage<L, TC * T> kys;
  kys.load(m, a.s);
  const int col = threadIdx.x % TC, t = threadIdx.x / TC;
  const Col cg = y_column<TC>(a.s, col);
  const bool valid = cg.valid;
  const int zb = cg.zb;
  const long S_ = cg.stride;
  const long base = cg.base;
  const gptr<float2> u0 = G(m.su[0]), u1 = G(m.su[1]), u2 = G(m.su[2]);
  const gptr<float2> g0 = G(m.sg[0]), g1 = G(m.sg[1]);
  float2 VX[S::R1][S::B], VY[S::R1][S::B], VZ[S::R1][S::B];
  in_load<L, T>(t, [&](int n) { return ld(u0, base + n * S_); }, VX);  // see k_ypass
  const float kzd = G(m.ktab)[2 * a.s.nx + 2 * a.s.ny + a.s.nzh + zb];
  tws.put(tw);
  kys.put(kyt);
  __syncthreads();
  // issued after the barrier (see k_ygrad): the first transform overlaps these loads
  in_load<L, T>(t, [&](int n) { return ld(u1, base + n * S_); }, VY);
  in_load<L, T>(t, [&](int n) { return ld(u2, base + n * S_); }, VZ);
  TileAt<TC> at{tile, col};
  float2 X[S::R2][S::A], XY[S::R2][S::A];
  in_compute<L, T, -1>(t, VX, at, tw);
  __syncthreads();
  in_finish<L, T, -1>(t, at, X);
  y_store<L, T>(t, valid, g0, base, S_, X, 1.f);
  __syncthreads();
  in_compute<L, T, -1>(t, VY, at, tw);
  __syncthreads();
  in_finish<L, T, -1>(t, at, XY);
  __syncthreads();
  in_compute<L, T, -1>(t, VZ, at, tw);
  __syncthreads();
  in_finish<L, T, -1>(t, at, X);
  // W = ky_d(k) FFT_y(U_y) + kz_d FFT_y(U_z), output bin k = k2 + B k1
#pragma unroll
  for (int r = 0; r < S::R2; ++r) {
    const int k2 = cmin(t + T * r, S::B - 1);
#pragma unroll
    for (int k1 = 0; k1 < S::A; ++k1) {
      const float f = kyt[k2 + S::B * k1];
      X[r][k1] = make_float2(f * XY[r][k1].x + kzd * X[r][k1].x, f * XY[r][k1].y + kzd * X[r][k1].y);
    }
  }
  y_store<L, T>(t, valid, g1, base, S_, X, 1.f);
}

// ================================================================ x passes
// Block = (y, column tile of TC bins); element (kx, y, zb) at kx * ps + y * ms + zb.
// Per column (y, zb): kappa / N at x-bin kx is sinc(c_ref dt / 2 |k|) / N with
// |k|^2 = kx^2 + ky^2 + kz^2.
struct XCol {
  const float2* kxt;  // LDS: (k_x, derivative k_x) per x bin
  float kyz2, h2, inv_n;
  __device__ __forceinline__ float kappa(int kxi) const {
    const float kx = kxt[kxi].x;
    return sinc_q(h2 * (kx * kx + kyz2)) * inv_n;
  }
  __device__ __forceinline__ float kxd(int kxi) const { return kxt[kxi].y; }
};

__device__ __forceinline__ XCol x_col(const MemberDev& m, const Shape& s, int y, int zb,
                                      const float2* kxt) {
  const gptr<const float> kt = G(m.ktab);
  const gptr<const float> ky_full = kt + 2 * s.nx;
  const gptr<const float> kz_full = ky_full + 2 * s.ny;
  XCol c;
  c.kxt = kxt;
  const float kyf = ky_full[y], kzf = kz_full[zb];
  c.kyz2 = kyf * kyf + kzf * kzf;
  c.h2 = m.half_cdt * m.half_cdt;
  c.inv_n = m.inv_n;
  return c;
}

// the x wavenumber pair table (k_x, derivative k_x) staged like the twiddles
template <int L, int NT>
struct KxStage {
  static constexpr int Q = (L + NT - 1) / NT;
  float2 v[Q];
  __device__ __forceinline__ void load(const MemberDev& m) {
    const gptr<const float> kt = G(m.ktab);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i0 = threadIdx.x + q * NT;
      const int i = i0 < L ? i0 : L - 1;
      v[q] = make_float2(kt[i], kt[L + i]);
    }
  }
  __device__ __forceinline__ void put(float2* kxt) const {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i = threadIdx.x + q * NT;
      if (i < L) kxt[i] = v[q];
    }
  }
};

// Waves per SIMD requested for k_xgrad: four (128 VGPRs) where the compiler fits the
// kernel without spilling (the default tile width), else no constraint. At L = 256 four
// waves spill four VGPRs to scratch; three waves without the spill run 1.5 % faster.
constexpr int xgrad_waves(int L, int TV) {
  return (TV == 0 && L < 256 && L != 144 && L != 200 && L != 240) ? 4 : 1;
}

// gradient: sg[0] = IFFT_x(kx_d i kappa / N FFT_x(sp)), sg[1] = Q = IFFT_x(i kappa / N
// FFT_x(sp)); the y and z factors (constant along x) are applied by k_ygrad.
template <int L, int TV = 0>
__global__ void __launch_bounds__((Plan<L, TV>::TC * Plan<L, TV>::T))
    __attribute__((amdgpu_waves_per_eu(xgrad_waves(L, TV))))
    k_xgrad(StepArgs a, const float2* __restrict__ twg) {
  using P = Plan<L, TV>;
  constexpr int T = P::T;
  using S = Split<L, T>;
  constexpr int TC = P::TC;
  __shared__ float2 tile[L * TC];
  __shared__ float2 tw[L];
  __shared__ float2 kxt[L];
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  // the first pass of a step advances the member's step clock and stores the step's
  // scalars (read by k_zp, which runs after this launch has completed): one extra block
  // per member (the first in x, dispatched at once) does only that, so that its chain of
  // dependent loads delays no column tile; column tiles are blocks 1, 2, ...
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
  const int col = threadIdx.x % TC, t = threadIdx.x / TC;
  const Col cg = x_column<TC>(a.s, col, (int)blockIdx.x - 1);
  const bool valid = cg.valid;
  const int y = cg.plane, zb = cg.zb;
  const long C = cg.stride;
  const long base = cg.base;
  float2 V[S::R1][S::B];
  const gptr<const float2> src = G(m.sp);
  in_load<L, T>(t, [&](int n) { return ld(src, base + n * C); }, V);  // see k_ypass
  tws.put(tw);
  kxs.put(kxt);
  __syncthreads();
  TileAt<TC> at{tile, col};
  in_compute<L, T, -1>(t, V, at, tw);
  __syncthreads();
  float2 X[S::R2][S::A];
  in_finish<L, T, -1>(t, at, X);
  const XCol c = x_col(m, a.s, y, zb, kxt);
  float2 Z[S::R2][S::A];
#pragma unroll
  for (int r = 0; r < S::R2; ++r) {
    const int k2 = t + T * r;
#pragma unroll
    for (int k1 = 0; k1 < S::A; ++k1) {
      const int kxi = (k2 < S::B) ? k2 + S::B * k1 : 0;
      const float kp = c.kappa(kxi);
      X[r][k1] = make_float2(-kp * X[r][k1].y, kp * X[r][k1].x);  // i kappa / N X
      const float kd = c.kxd(kxi);
      Z[r][k1] = make_float2(kd * X[r][k1].x, kd * X[r][k1].y);
    }
  }
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    __syncthreads();
    out_compute<L, T, +1>(t, pass == 0 ? Z : X, at, tw);
    __syncthreads();
    out_finish<L, T, +1>(t, at, V);
    if (valid) {
      if (pass == 0) {
        const gptr<float2> dst = G(m.sg[0]);
#pragma unroll
        for (int r = 0; r < S::R1; ++r) {
          const int m2 = t + T * r;
          if (m2 < S::A) {
#pragma unroll
            for (int j = 0; j < S::B; ++j) dst[base + (long)(m2 + S::A * j) * C] = V[r][j];
          }
        }
      } else {
        const gptr<float2> dq = G(m.sg[1]);  // Q; k_ygrad applies ky_d and kz_d
#pragma unroll
        for (int r = 0; r < S::R1; ++r) {
          const int m2 = t + T * r;
          if (m2 < S::A) {
#pragma unroll
            for (int j = 0; j < S::B; ++j) dq[base + (long)(m2 + S::A * j) * C] = V[r][j];
          }
        }
      }
    }
  }
}

// divergence: sp = IFFT_x( i kappa / N (kx_d FFT_x(sg[0]) + FFT_x(sg[1])) ), with sg[1] the
// y/z divergence terms already combined by k_ydiv
template <int L, int TV = 0>
__global__ void __launch_bounds__((Plan<L, TV>::TC * Plan<L, TV>::T))
    k_xdiv(StepArgs a, const float2* __restrict__ twg) {
  using P = Plan<L, TV>;
  constexpr int T = P::T;
  using S = Split<L, T>;
  constexpr int TC = P::TC;
  __shared__ float2 tile[L * TC];
  __shared__ float2 tw[L];
  __shared__ float2 kxt[L];
  TwStage<L, TC * T> tws;
  tws.load(twg);
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  KxStage<L, TC * T> kxs;
  kxs.load(m);
  const int col = threadIdx.x % TC, t = threadIdx.x / TC;
  const Col cg = x_column<TC>(a.s, col);
  const bool valid = cg.valid;
  const int y = cg.plane, zb = cg.zb;
  const long C = cg.stride;
  const long base = cg.base;
  float2 V[S::R1][S::B], W[S::R1][S::B];
  const gptr<const float2> ux = G(m.sg[0]), uw = G(m.sg[1]);
  in_load<L, T>(t, [&](int n) { return ld(ux, base + n * C); }, V);  // see k_ypass
  const XCol c = x_col(m, a.s, y, zb, kxt);
  TileAt<TC> at{tile, col};
  tws.put(tw);
  kxs.put(kxt);
  __syncthreads();
  // issued after the barrier: the first transform then waits for U_x only, and this
  // load overlaps it
  in_load<L, T>(t, [&](int n) { return ld(uw, base + n * C); }, W);
  in_compute<L, T, -1>(t, V, at, tw);
  __syncthreads();
  float2 Xx[S::R2][S::A];
  in_finish<L, T, -1>(t, at, Xx);
  __syncthreads();
  in_compute<L, T, -1>(t, W, at, tw);
  __syncthreads();
  float2 X[S::R2][S::A];
  in_finish<L, T, -1>(t, at, X);
#pragma unroll
  for (int r = 0; r < S::R2; ++r) {
    const int k2 = t + T * r;
#pragma unroll
    for (int k1 = 0; k1 < S::A; ++k1) {
      const int kxi = (k2 < S::B) ? k2 + S::B * k1 : 0;
      const float kd = c.kxd(kxi);
      const float kp = c.kappa(kxi);
      const float2 sum = make_float2(kd * Xx[r][k1].x + X[r][k1].x, kd * Xx[r][k1].y + X[r][k1].y);
      X[r][k1] = make_float2(-kp * sum.y, kp * sum.x);
    }
  }
  __syncthreads();
  out_compute<L, T, +1>(t, X, at, tw);
  __syncthreads();
  out_finish<L, T, +1>(t, at, V);
  if (valid) {
    const gptr<float2> dst = G(m.sp);
#pragma unroll
    for (int r = 0; r < S::R1; ++r) {
      const int m2 = t + T * r;
      if (m2 < S::A) {
#pragma unroll
        for (int j = 0; j < S::B; ++j) dst[base + (long)(m2 + S::A * j) * C] = V[r][j];
      }
    }
  }
}

// ================================================================ z passes
// Real length nz = 2M along the contiguous axis, computed as a length-M complex FFT of
// z[m] = x[2m] + i x[2m+1] with the even/odd split (rt[k] = exp(2 pi i k / nz)).
// Per line, LDS holds a natural-order buffer (M + 1 bins, later the real line as M
// float2) and an exchange buffer; all HBM traffic of the pass is contiguous: the
// spectrum rows, and the real volumes as one float2 stream per block.

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

template <int M, int ZV>
struct ZLds {
  static constexpr int NA = ZPlan<M, ZV>::LPB * ZPlan<M, ZV>::LSA;
  static constexpr int NE = ZPlan<M, ZV>::LPB * ZPlan<M, ZV>::LSE;
  static constexpr bool ALIAS = ZPlan<M, ZV>::ALIAS;
  float2 a[ALIAS ? cmax(NA, NE) : NA];
  float2 ex[ALIAS ? 1 : NE];
  __device__ __forceinline__ float2* e() {
    if constexpr (ALIAS) return a;
    else return ex;
  }
  float2 tw[M];
  float2 rt[M + 1];
  float2 sz[M];  // z sponge profile as pairs (z = 2 mm, 2 mm + 1)
  float sxy[ZPlan<M, ZV>::LPB];
  int lx[ZPlan<M, ZV>::LPB];
  int ly[ZPlan<M, ZV>::LPB];
  long row[ZPlan<M, ZV>::LPB];  // offset of each line's spectrum row
  int sptr[ZPlan<M, ZV>::LPB + 1];  // source CSR bounds of the block's lines
};

// per-thread count of the block's contiguous real chunk (float2) and of its row bins
template <int M, int ZV>
struct ZCount {
  static constexpr int EPT = (ZPlan<M, ZV>::LPB * M + ZPlan<M, ZV>::THREADS - 1) / ZPlan<M, ZV>::THREADS;
};

// (x, y) of line line0 + l of a block. Lines of a block are consecutive, so when the
// block holds at most ny lines they cross at most one x plane: one compare instead of a
// per-thread integer division. x0, y0 (the block's first line) are block-uniform.
// NARROW: the caller has established LPB <= ny (block-uniform), so the test is not
// repeated per element.
template <int LPB, bool NARROW>
__device__ __forceinline__ void line_xy_(int x0, int y0, int l, const Shape& sh, int& x, int& y) {
  y = y0 + l;
  x = x0;
  if constexpr (NARROW) {
    const bool w = y >= sh.ny;
    y -= w ? sh.ny : 0;
    x += w ? 1 : 0;
  } else {
    x += y / sh.ny;
    y -= (y / sh.ny) * sh.ny;
  }
}
template <int LPB>
__device__ __forceinline__ void line_xy(int x0, int y0, int l, const Shape& sh, int& x, int& y) {
  if (LPB <= sh.ny) line_xy_<LPB, true>(x0, y0, l, sh, x, y);
  else line_xy_<LPB, false>(x0, y0, l, sh, x, y);
}

// n / D for 0 <= n < NMAX as one full-rate 24-bit multiply and a shift (the compiler's
// division by a constant uses a quarter-rate 32-bit high multiply). The multiplier is
// checked exhaustively at compile time; when no exact one fits, plain division is used.
template <int D, int NMAX>
struct Div24 {
  static constexpr unsigned S = 23;
  static constexpr unsigned MG = (1u << S) / D + 1;
  static constexpr bool exact() {
    if (MG >= (1u << 24) || NMAX >= (1 << 24)) return false;
    for (unsigned n = 0; n < (unsigned)NMAX; ++n) {
      const unsigned long long p = (unsigned long long)n * MG;
      if (p >= (1ull << 32) || (unsigned)(p >> S) != n / D) return false;
    }
    return true;
  }
  static constexpr bool OK = exact();
  __device__ __forceinline__ static int div(int n) {
    if constexpr (OK) return (int)(__umul24((unsigned)n, MG) >> S);
    else return n / D;
  }
};

// Element j of a thread in a z-pass block's chunk of LPB lines x M float2: chunk index
// i = threadIdx.x + j THREADS, line l = i / M, position mm = i % M; i < nl M exactly
// when l < nl. When THREADS is a multiple of M, mm is the same for every j and l steps
// by THREADS / M; when M is a multiple of THREADS, l is a compile-time constant per j.
// Then no element needs a division, and where l is wave-uniform (THREADS = M, or M a
// multiple of THREADS) the line-dependent parts of every address are scalar. Other
// shapes divide with Div24.
template <int M, int ZV>
struct ZElem {
  static constexpr int T = ZPlan<M, ZV>::THREADS;
  static constexpr bool ROWS = T % M == 0;
  static constexpr bool SPLIT = !ROWS && M % T == 0;
  __device__ __forceinline__ static int line(int j) {
    if constexpr (ROWS) {
      // a wave's 64 lanes share threadIdx.x / M when M is a multiple of 64: say so
      if constexpr (M % 64 == 0) return j * (T / M) + __builtin_amdgcn_readfirstlane((int)threadIdx.x / M);
      else return j * (T / M) + (int)threadIdx.x / M;
    } else if constexpr (SPLIT) {
      return j / (M / T);
    } else {
      return Div24<M, ZCount<M, ZV>::EPT * T>::div((int)threadIdx.x + j * T);
    }
  }
  __device__ __forceinline__ static int pos(int j, int l) {
    if constexpr (ROWS) return (int)threadIdx.x % M;
    else if constexpr (SPLIT) return (j % (M / T)) * T + (int)threadIdx.x;
    else return (int)threadIdx.x + j * T - l * M;
  }
};

// Tables a block stages into LDS (twiddles, real-split factors, z sponge, per-line
// sponge product and spectrum-row offsets). They are loaded into registers first and
// written to LDS after the block's HBM loads are issued: a wait for a load also waits
// for every load issued before it, so anything consumed early must be loaded early.
template <int M, int ZV>
struct ZStage {
  static constexpr int THREADS = ZPlan<M, ZV>::THREADS;
  static constexpr int TPT = (M + 1 + THREADS - 1) / THREADS;
  float2 tw[TPT], rt[TPT], sz[TPT];
  float sx, sy;
};

// damping table `tab`: sz[nz] sx[nx] sy[ny] (the member's sponge, or a table of the
// staggered scheme)
template <int M, int ZV>
__device__ __forceinline__ void z_stage_load(const float2* twg, const float2* rtg, const float* tab,
                                             const Shape& sh, long line0, int nl, ZStage<M, ZV>& st) {
  // every load is unconditional (indices clamped into range): the wait-count pass can
  // then wait for exactly the loads a consumer needs instead of for all of them
  constexpr int TH = ZStage<M, ZV>::THREADS;
  constexpr int A = ZPlan<M, ZV>::A;
  const gptr<const float2> szg = G(reinterpret_cast<const float2*>(tab));
#pragma unroll
  for (int q = 0; q < ZStage<M, ZV>::TPT; ++q) {
    const int i = threadIdx.x + q * TH;
    const int im = i < M ? i : M - 1;
    st.tw[q] = twg[(im % A) * (im / A)];  // transposed table (see in_compute)
    st.sz[q] = szg[im];
    st.rt[q] = rtg[i <= M ? i : M];
  }
  const int x0 = (int)line0 / sh.ny, y0 = (int)line0 - x0 * sh.ny;  // block-uniform
  int x, y;
  line_xy<ZPlan<M, ZV>::LPB>(x0, y0, (int)threadIdx.x < nl ? (int)threadIdx.x : nl - 1, sh, x, y);
  const gptr<const float> sxg = G(tab) + 2 * M;
  st.sx = sxg[x];
  st.sy = sxg[sh.nx + y];
}

template <int M, int ZV>
__device__ __forceinline__ void z_stage_load(const float2* twg, const float2* rtg, const MemberDev& m,
                                             const Shape& sh, long line0, int nl, ZStage<M, ZV>& st) {
  z_stage_load<M, ZV>(twg, rtg, m.sponge, sh, line0, nl, st);
}

template <int M, int ZV>
__device__ __forceinline__ void z_stage_put(const ZStage<M, ZV>& st, const Shape& sh, long line0, int nl,
                                            ZLds<M, ZV>& s) {
  constexpr int TH = ZStage<M, ZV>::THREADS;
#pragma unroll
  for (int q = 0; q < ZStage<M, ZV>::TPT; ++q) {
    const int i = threadIdx.x + q * TH;
    if (i < M) {
      s.tw[i] = st.tw[q];
      s.sz[i] = st.sz[q];
    }
    if (i <= M) s.rt[i] = st.rt[q];
  }
  if (threadIdx.x < nl) {
    const int x0 = (int)line0 / sh.ny, y0 = (int)line0 - x0 * sh.ny;
    int x, y;
    line_xy<ZPlan<M, ZV>::LPB>(x0, y0, threadIdx.x, sh, x, y);
    s.sxy[threadIdx.x] = st.sx * st.sy;
    s.lx[threadIdx.x] = x;
    s.ly[threadIdx.x] = y;
    s.row[threadIdx.x] = (long)x * sh.ps + (long)y * sh.ms;
  }
}

// issue the loads of the block's spectra into registers: bins k < M of each line from its
// row (thread-to-element mapping of the real volumes), and the Nyquist bin k = M of line
// line0 + threadIdx.x from the side plane
template <int M, int ZV, bool NARROW>
__device__ __forceinline__ void z_spec_load_(gptr<const float2> spec, long line0, int nl,
                                             const Shape& sh, float2 (&r)[ZCount<M, ZV>::EPT],
                                             float2& rn) {
  using E = ZElem<M, ZV>;
  const int x0 = (int)line0 / sh.ny, y0 = (int)line0 - x0 * sh.ny;  // block-uniform
  // row of line (x, y) = x ps + y ms = line ms + 16 x (ps = ny ms + 16): offsets from the
  // block's first row fit 32 bits and need no per-element 64-bit multiply
  const gptr<const float2> base = spec + ((long)line0 * sh.ms + 16L * x0);
#pragma unroll
  for (int j = 0; j < ZCount<M, ZV>::EPT; ++j) {
    const int l0 = E::line(j), k = E::pos(j, l0);
    const int l = l0 < nl ? l0 : nl - 1;  // unconditional load (see z_stage_load)
    int x, y;
    line_xy_<ZPlan<M, ZV>::LPB, NARROW>(x0, y0, l, sh, x, y);
    // 24-bit multiply and unsigned sums: the offset stays a full-rate 32-bit computation
    const gptr<const float2> row = base + (__umul24((unsigned)l, (unsigned)sh.ms) + 16u * (unsigned)(x - x0));
    r[j] = row[k];
  }
  rn = spec[sh.so + line0 + ((int)threadIdx.x < nl ? (int)threadIdx.x : nl - 1)];
}

// the plane test of line_xy is block-uniform: taken once here, not per unrolled element
// (per element it left every load in a basic block of its own behind a scalar branch)
template <int M, int ZV>
__device__ __forceinline__ void z_spec_load(gptr<const float2> spec, long line0, int nl,
                                            const Shape& sh, float2 (&r)[ZCount<M, ZV>::EPT],
                                            float2& rn) {
  if (ZPlan<M, ZV>::LPB <= sh.ny) z_spec_load_<M, ZV, true>(spec, line0, nl, sh, r, rn);
  else z_spec_load_<M, ZV, false>(spec, line0, nl, sh, r, rn);
}

template <int M, int ZV>
__device__ __forceinline__ void z_spec_put(int nl, const float2 (&r)[ZCount<M, ZV>::EPT], float2 rn,
                                           ZLds<M, ZV>& s) {
  using E = ZElem<M, ZV>;
#pragma unroll
  for (int j = 0; j < ZCount<M, ZV>::EPT; ++j) {
    const int l = E::line(j), k = E::pos(j, l);
    if (l < nl) s.a[l * ZPlan<M, ZV>::LSA + k] = r[j];
  }
  if ((int)threadIdx.x < nl) s.a[threadIdx.x * ZPlan<M, ZV>::LSA + M] = rn;
}

// issue the loads of a real volume's chunk (float2 stream) into registers
template <int M, int ZV>
__device__ __forceinline__ void z_vol_load(gptr<const float2> v, int nl,
                                           float2 (&r)[ZCount<M, ZV>::EPT]) {
  using E = ZElem<M, ZV>;
#pragma unroll
  for (int j = 0; j < ZCount<M, ZV>::EPT; ++j) {
    const int l0 = E::line(j), mm = E::pos(j, l0);
    const int l = l0 < nl ? l0 : nl - 1;  // unconditional load (see z_stage_load)
    r[j] = (v + l * M)[mm];
  }
}

// Second stage of a z-pass line transform: exchange buffer le -> natural order in la.
// Shapes with twice as many threads per line as stage-2 DFTs share each DFT between two
// threads (PairFinish), the others run the plain finish.
template <int M, int ZV, int DIR>
__device__ __forceinline__ void z_finish_store(int t, bool lvalid, float2* le, float2* la) {
  using Z = ZPlan<M, ZV>;
  using S = Split<M, Z::ZT, Z::A>;
  if constexpr (PairFinish<M, Z::ZT, Z::A>::OK) {
    constexpr int H = PairFinish<M, Z::ZT, Z::A>::H;
    float2 Y[H];
    if (lvalid) in_finish_pair<M, Z::ZT, DIR, Z::A>(t, LinePadAt<S::A>{le}, Y);
    if constexpr (Z::ALIAS) __syncthreads();  // e consumed before a (the same array) is written
    if (lvalid) {
#pragma unroll
      for (int m = 0; m < H; ++m) la[t + Z::ZT * m] = Y[m];
    }
  } else {
    float2 X[S::R2][S::A];
    if (lvalid) in_finish<M, Z::ZT, DIR, Z::A>(t, LinePadAt<S::A>{le}, X);
    if constexpr (Z::ALIAS) __syncthreads();  // e consumed before a (the same array) is written
    if (lvalid) {
#pragma unroll
      for (int r = 0; r < S::R2; ++r) {
        const int k2 = t + Z::ZT * r;
        if (k2 < S::B) {
#pragma unroll
          for (int k1 = 0; k1 < S::A; ++k1) la[k2 + S::B * k1] = X[r][k1];
        }
      }
    }
  }
}

// C2R of every staged line: natural buffer a -> real line (as M float2) back in a.
template <int M, int ZV>
__device__ __forceinline__ void z_c2r(int ll, int t, bool lvalid, ZLds<M, ZV>& s) {
  using Z = ZPlan<M, ZV>;
  using S = Split<M, Z::ZT, Z::A>;
  float2* la = s.a + ll * Z::LSA;
  float2* le = s.e() + ll * Z::LSE;
  float2 V[S::R1][S::B];
  in_load<M, Z::ZT, Z::A>(t, [&](int n) { return c2r_pre<M>(la, n, s.rt); }, V);
  __syncthreads();  // a consumed
  if (lvalid) in_compute<M, Z::ZT, +1, true, Z::A>(t, V, LinePadAt<S::A>{le}, s.tw);
  __syncthreads();
  z_finish_store<M, ZV, +1>(t, lvalid, le, la);
  __syncthreads();
}

// R2C of the real line in a (M float2) -> spectrum row in HBM (bins k < M) and the
// line's Nyquist entry of the side plane (dst_nyq).
template <int M, int ZV>
__device__ __forceinline__ void z_r2c(int ll, int t, bool lvalid, ZLds<M, ZV>& s,
                                      gptr<float2> dst_row, gptr<float2> dst_nyq) {
  using Z = ZPlan<M, ZV>;
  using S = Split<M, Z::ZT, Z::A>;
  float2* la = s.a + ll * Z::LSA;
  float2* le = s.e() + ll * Z::LSE;
  float2 V[S::R1][S::B];
  in_load<M, Z::ZT, Z::A>(t, [&](int n) { return la[n]; }, V);
  if constexpr (Z::ALIAS) __syncthreads();  // a consumed before e (the same array) is written
  if (lvalid) in_compute<M, Z::ZT, -1, true, Z::A>(t, V, LinePadAt<S::A>{le}, s.tw);
  __syncthreads();  // e written, a consumed
  z_finish_store<M, ZV, -1>(t, lvalid, le, la);
  __syncthreads();
  if (lvalid) {
    for (int k = t; k < M; k += Z::ZT) dst_row[k] = r2c_post<M>(la, k, s.rt);
    if (t == 0) *dst_nyq = r2c_post<M>(la, M, s.rt);
  }
}

// Waves per SIMD requested for k_zu: four (128 VGPRs) for layouts 0 and 2 (their LDS
// allows four per SIMD), except at the lengths where the compiler then spills to
// scratch (for layout 2 with one-wave blocks, whose shared line buffer leaves the LDS
// room for four: 150, 160, 192, 288, 300); else no constraint.
constexpr int zu_waves(int M, int ZV) {
  return (zlayout(ZV) != 1 && (ZV >> 2) == 0 && M != 72 && M != 80 && M != 96 && M != 100 &&
          M != 144 && M != 200 &&
          !(zlayout(ZV) == 2 && (M == 120 || M == 150 || M == 160 || M == 192 || M == 288 || M == 300)))
             ? 4
             : 1;
}

// velocity update, one axis per block. The velocity state is kept as its z-spectrum
// (su), the form the divergence needs next, so a component is read and written once
// per step as a spectrum instead of as a real volume plus its spectrum: per axis
// 24 H bytes instead of 16 H + 8 P. The block C2Rs the gradient spectrum, keeps the
// real gradient in registers, C2Rs the velocity spectrum, updates
// u = ((u - dt/rho g) absorb) sponge in LDS and R2Cs it back into su.
// Block b covers line group ngroups - 1 - g and axis ax with b = 24 * (g / 8) + 8 * ax + (g % 8):
// blocks are dealt to the 8 XCDs round-robin, so the three axis blocks of a line group
// run back to back on one XCD and its L2 serves the second and third reads of dt/rho
// and absorb.
template <int M, int ZV>
__global__ void __launch_bounds__((ZPlan<M, ZV>::THREADS))
    __attribute__((amdgpu_waves_per_eu(zu_waves(M, ZV))))
    k_zu(StepArgs a, const float2* __restrict__ twg, const float2* __restrict__ rtg) {
  using Z = ZPlan<M, ZV>;
  using CN = ZCount<M, ZV>;
  __shared__ ZLds<M, ZV> s;
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  const int nlines = a.s.nx * a.s.ny;
  const int ngroups = (nlines + Z::LPB - 1) / Z::LPB;
  const int b = blockIdx.x;
  const int g0 = (b / 24) * 8 + (b % 8);
  const int ax = (b / 8) % 3;
  if (g0 >= ngroups) return;
  // line groups in reverse: the y pass before this one wrote the spectra in increasing x,
  // so the lines it wrote last (still in the last-level cache) are read first
  const int g = ngroups - 1 - g0;
  const long line0 = (long)g * Z::LPB;
  const int nl = cmin(Z::LPB, nlines - (int)line0);
  const int ll = threadIdx.x / Z::ZT, t = threadIdx.x % Z::ZT;
  const bool lvalid = ll < nl;
  const long e0 = line0 * M;  // first float2 of the block's chunk of a real volume
  const gptr<float2> gspec = G(m.sg[ax]);
  const gptr<float2> uspec = G(m.su[ax]);
  const float inz = m.inv_nz;
  // loads in the order they are consumed: staged tables, gradient spectrum, velocity
  // spectrum; the maps after the first transform (their registers are not held through
  // it, which keeps the kernel at four waves per SIMD)
  ZStage<M, ZV> stg;
  z_stage_load<M, ZV>(twg, rtg, m, a.s, line0, nl, stg);
  float2 sr[CN::EPT], ur[CN::EPT], dto[CN::EPT], ab[CN::EPT], srn, urn;
  z_spec_load<M, ZV>(gspec, line0, nl, a.s, sr, srn);
  z_spec_load<M, ZV>(uspec, line0, nl, a.s, ur, urn);
  z_stage_put<M, ZV>(stg, a.s, line0, nl, s);
  z_spec_put<M, ZV>(nl, sr, srn, s);
  __syncthreads();
  z_c2r<M, ZV>(ll, t, lvalid, s);
  z_vol_load<M, ZV>(G(reinterpret_cast<const float2*>(m.dto_rho)) + e0, nl, dto);
  z_vol_load<M, ZV>(G(reinterpret_cast<const float2*>(m.absorb)) + e0, nl, ab);
  float2 gg[CN::EPT];
#pragma unroll
  for (int j = 0; j < CN::EPT; ++j) {
    const int l0 = ZElem<M, ZV>::line(j), mm = ZElem<M, ZV>::pos(j, l0);
    const int l = l0 < nl ? l0 : nl - 1;
    gg[j] = s.a[l * Z::LSA + mm];
  }
  __syncthreads();
  z_spec_put<M, ZV>(nl, ur, urn, s);
  __syncthreads();
  z_c2r<M, ZV>(ll, t, lvalid, s);
#pragma unroll
  for (int j = 0; j < CN::EPT; ++j) {
    const int l = ZElem<M, ZV>::line(j), mm = ZElem<M, ZV>::pos(j, l);
    if (l < nl) {
      const float sxy = s.sxy[l];
      const float2 s2 = s.sz[mm];
      float2 uv = s.a[l * Z::LSA + mm];
      uv.x *= inz;
      uv.y *= inz;
      uv.x -= dto[j].x * gg[j].x;
      uv.y -= dto[j].y * gg[j].y;
      uv.x *= ab[j].x;
      uv.y *= ab[j].y;
      uv.x *= sxy * s2.x;
      uv.y *= sxy * s2.y;
      s.a[l * Z::LSA + mm] = uv;
    }
  }
  __syncthreads();
  z_r2c<M, ZV>(ll, t, lvalid, s, uspec + s.row[lvalid ? ll : 0], uspec + a.s.so + line0 + (lvalid ? ll : 0));
}

// Conversion of one velocity component between its state form and a real volume, for
// reading and writing the state: dir > 0: R2C of the real volume held in sg[ax]'s memory
// into su[ax]; dir < 0: C2R of su[ax], divided by nz, into sg[ax]'s memory as a real
// volume (z contiguous, the layout of p). sg is scratch between steps.
template <int M, int ZV>
__global__ void __launch_bounds__((ZPlan<M, ZV>::THREADS))
    k_zconv(StepArgs a, const float2* __restrict__ twg, const float2* __restrict__ rtg, int ax, int dir) {
  using Z = ZPlan<M, ZV>;
  using CN = ZCount<M, ZV>;
  __shared__ ZLds<M, ZV> s;
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  const int nlines = a.s.nx * a.s.ny;
  const long line0 = (long)blockIdx.x * Z::LPB;
  const int nl = cmin(Z::LPB, nlines - (int)line0);
  const int ll = threadIdx.x / Z::ZT, t = threadIdx.x % Z::ZT;
  const bool lvalid = ll < nl;
  const gptr<float2> real = G(m.sg[ax]) + line0 * M;
  const gptr<float2> uspec = G(m.su[ax]);
  ZStage<M, ZV> stg;
  z_stage_load<M, ZV>(twg, rtg, m, a.s, line0, nl, stg);
  if (dir > 0) {
    float2 vv[CN::EPT];
    z_vol_load<M, ZV>(real, nl, vv);
    z_stage_put<M, ZV>(stg, a.s, line0, nl, s);
#pragma unroll
    for (int j = 0; j < CN::EPT; ++j) {
      const int l = ZElem<M, ZV>::line(j), mm = ZElem<M, ZV>::pos(j, l);
      if (l < nl) s.a[l * Z::LSA + mm] = vv[j];
    }
    __syncthreads();
    z_r2c<M, ZV>(ll, t, lvalid, s, uspec + s.row[lvalid ? ll : 0], uspec + a.s.so + line0 + (lvalid ? ll : 0));
  } else {
    const float inz = m.inv_nz;
    float2 sr[CN::EPT], srn;
    z_spec_load<M, ZV>(uspec, line0, nl, a.s, sr, srn);
    z_stage_put<M, ZV>(stg, a.s, line0, nl, s);
    z_spec_put<M, ZV>(nl, sr, srn, s);
    __syncthreads();
    z_c2r<M, ZV>(ll, t, lvalid, s);
#pragma unroll
    for (int j = 0; j < CN::EPT; ++j) {
      const int l = ZElem<M, ZV>::line(j), mm = ZElem<M, ZV>::pos(j, l);
      if (l < nl) {
        const float2 v = s.a[l * Z::LSA + mm];
        (real + l * M)[mm] = make_float2(v.x * inz, v.y * inz);
      }
    }
  }
}

// pressure update (+ Westervelt), source, convergence peak, record; or (update == 0) a
// plain R2C of the current p. After the C2R every element is finished in one pass:
// update, the source voxels of its line (values computed during the prefetch and
// parked in LDS), the store of p, the convergence peak and the record sums.
// NL = false is launched when no member of the batch is nonlinear: it holds no
// registers for the beta map, which lets one more wave per SIMD in.
template <int M, int ZV, bool NL>
__global__ void __launch_bounds__((ZPlan<M, ZV>::THREADS))
    k_zp(StepArgs a, const float2* __restrict__ twg, const float2* __restrict__ rtg, int update) {
  using Z = ZPlan<M, ZV>;
  using CN = ZCount<M, ZV>;
  constexpr int kSrcPT = 1;                       // prefetched source entries per thread
  constexpr int kSrcCap = kSrcPT * Z::THREADS;    // entries a block keeps in LDS
  __shared__ ZLds<M, ZV> s;
  __shared__ int src_zl[kSrcCap];
  __shared__ float src_val[kSrcCap];
  const MemberDev& m = a.mem[a.ids[blockIdx.z]];
  const StepMember sm = update ? member_step(m) : StepMember{};
  const int nlines = a.s.nx * a.s.ny;
  const long line0 = (long)(gridDim.x - 1 - blockIdx.x) * Z::LPB;  // reverse order, as in k_zu
  const int nl = cmin(Z::LPB, nlines - (int)line0);
  const int ll = threadIdx.x / Z::ZT, t = threadIdx.x % Z::ZT;
  const bool lvalid = ll < nl;
  const long e0 = line0 * M;
  const gptr<float2> p = G(reinterpret_cast<float2*>(m.p)) + e0;
  float2 pp[CN::EPT];
  if (update) {
    // loads in the order they are consumed: staged tables, spectrum, source list (all
    // before the C2R), then the volumes (after it: their latency overlaps the transform)
    // the block's source bounds first: they are block-uniform (scalar) loads, and the
    // source entries' vector loads depend on them
    const gptr<const int> sptr = G(m.src_ptr);
    const int sbeg = sptr[line0], send = sptr[line0 + nl];
    ZStage<M, ZV> stg;
    z_stage_load<M, ZV>(twg, rtg, m, a.s, line0, nl, stg);
    float2 sr[CN::EPT], rc[CN::EPT], ab[CN::EPT], bt[NL ? CN::EPT : 1], srn;
    z_spec_load<M, ZV>(G(m.sp), line0, nl, a.s, sr, srn);
    const bool src_fit = send - sbeg <= kSrcCap;  // block-uniform
    int sz_pf[kSrcPT];
    float2 sw_pf[kSrcPT];
    int sp_l = 0;
    // unconditional loads (see z_stage_load); the source arrays hold at least one entry
#pragma unroll
    for (int q = 0; q < kSrcPT; ++q) {
      const int j = sbeg + threadIdx.x + q * Z::THREADS;
      const int jc = j < send ? j : (send > 0 ? send - 1 : 0);
      sz_pf[q] = G(m.src_z)[jc];
      sw_pf[q] = G(m.src_ab)[jc];
    }
    sp_l = sptr[line0 + ((int)threadIdx.x <= nl ? (int)threadIdx.x : nl)];
    z_vol_load<M, ZV>(p, nl, pp);
    z_vol_load<M, ZV>(G(reinterpret_cast<const float2*>(m.rhoc2dt)) + e0, nl, rc);
    z_vol_load<M, ZV>(G(reinterpret_cast<const float2*>(m.absorb)) + e0, nl, ab);
    if constexpr (NL) {
      if (m.beta2dt) z_vol_load<M, ZV>(G(reinterpret_cast<const float2*>(m.beta2dt)) + e0, nl, bt);
    }
    z_stage_put<M, ZV>(stg, a.s, line0, nl, s);
    if (threadIdx.x <= nl) s.sptr[threadIdx.x] = sp_l;
    // a block with as many lines as threads has one bound more than threads
    for (int i = (int)threadIdx.x + Z::THREADS; i <= nl; i += Z::THREADS) s.sptr[i] = sptr[line0 + i];
#pragma unroll
    for (int q = 0; q < kSrcPT; ++q) {
      const int j = sbeg + threadIdx.x + q * Z::THREADS;
      if (j < send) {
        src_zl[j - sbeg] = sz_pf[q];
        src_val[j - sbeg] = sm.amp * (sw_pf[q].x * sm.sin_t - sw_pf[q].y * sm.cos_t);
      }
    }
    z_spec_put<M, ZV>(nl, sr, srn, s);
    __syncthreads();
    z_c2r<M, ZV>(ll, t, lvalid, s);
    // descriptor fields used after the first store, read once here: a global store may
    // alias the descriptor table as far as the compiler knows, and would force re-reads
    // with vector loads inside the element loop
    int cv[6], rb[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      cv[q] = m.conv[q];
      rb[q] = m.rec[q];
    }
    const int nh = m.nh;
    const bool nonlin = NL && m.beta2dt != nullptr;
    gptr<float2> accp[kMaxHarmonics], compp[kMaxHarmonics];
#pragma unroll
    for (int h = 0; h < kMaxHarmonics; ++h) {
      accp[h] = G(m.acc[h]);
      compp[h] = G(m.comp[h]);
    }
    const gptr<float> pmaxp = G(m.pmax);
    unsigned* const peaksp = m.peaks;
    const int ry = rb[3] - rb[2], rz = rb[5] - rb[4];
    unsigned pk = 0u;
    // store of the finished element (l, mm) at chunk offset i, its convergence peak and,
    // on a record step, its phasor (Kahan) and p_max updates
    auto finish = [&](int i, int l, int mm, float2 pv) {
      const int x = s.lx[l], y = s.ly[l];
      p[i] = pv;
      if (x >= cv[0] && x < cv[1] && y >= cv[2] && y < cv[3]) {
        const int z0 = 2 * mm;
        if (z0 >= cv[4] && z0 < cv[5]) pk = max(pk, __float_as_uint(fabsf(pv.x)));
        if (z0 + 1 >= cv[4] && z0 + 1 < cv[5]) pk = max(pk, __float_as_uint(fabsf(pv.y)));
      }
      if (sm.record && x >= rb[0] && x < rb[1] && y >= rb[2] && y < rb[3]) {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int zz = 2 * mm + e;
          const float v = e ? pv.y : pv.x;
          if (zz >= rb[4] && zz < rb[5]) {
            const long ri = ((long)(x - rb[0]) * ry + (y - rb[2])) * rz + (zz - rb[4]);
#pragma unroll
            for (int h = 0; h < kMaxHarmonics; ++h) {
              if (h >= nh) break;
              const gptr<float2> acc = accp[h], comp = compp[h];
              const float2 eh = sm.eh[h];
              const float2 sacc = acc[ri];
              float2 cp = comp[ri];
              const float yx = v * eh.x - cp.x, yy = v * eh.y - cp.y;
              const float tx = sacc.x + yx, ty = sacc.y + yy;
              cp.x = (tx - sacc.x) - yx;
              cp.y = (ty - sacc.y) - yy;
              acc[ri] = make_float2(tx, ty);
              comp[ri] = cp;
            }
            pmaxp[ri] = fmaxf(pmaxp[ri], fabsf(v));
          }
        }
      }
    };
#pragma unroll
    for (int j = 0; j < CN::EPT; ++j) {
      const int l = ZElem<M, ZV>::line(j), mm = ZElem<M, ZV>::pos(j, l);
      if (l < nl) {
        const int i = l * M + mm;
        const float sxy = s.sxy[l];
        const float2 dv = s.a[l * Z::LSA + mm];
        float2 pv = pp[j];
        float2 c = rc[j];
        if constexpr (NL) {
          if (nonlin) {
            c.x += bt[j].x * pv.x;
            c.y += bt[j].y * pv.y;
          }
        }
        const float2 s2 = s.sz[mm];
        pv.x -= c.x * dv.x;
        pv.y -= c.y * dv.y;
        pv.x *= ab[j].x;
        pv.y *= ab[j].y;
        pv.x *= sxy * s2.x;
        pv.y *= sxy * s2.y;
        if (src_fit) {
          for (int e = s.sptr[l]; e < s.sptr[l + 1]; ++e) {
            const int zz = src_zl[e - sbeg];
            if ((zz >> 1) == mm) {
              if (zz & 1) pv.y += src_val[e - sbeg];
              else pv.x += src_val[e - sbeg];
            }
          }
          finish(i, l, mm, pv);
        }
        s.a[l * Z::LSA + mm] = pv;
      }
    }
    if (!src_fit) {
      // more source voxels than the LDS list holds: add them in a separate pass
      __syncthreads();
      const gptr<const int> srcz = G(m.src_z);
      const gptr<const float2> srcab = G(m.src_ab);
      for (int j = sbeg + threadIdx.x; j < send; j += Z::THREADS) {
        int l = 0;
        while (l + 1 < nl && s.sptr[l + 1] <= j) ++l;
        const float2 w = srcab[j];
        float* lf = reinterpret_cast<float*>(s.a + l * Z::LSA);
        lf[srcz[j]] += sm.amp * (w.x * sm.sin_t - w.y * sm.cos_t);
      }
      __syncthreads();
#pragma unroll 1
      for (int i = threadIdx.x; i < nl * M; i += Z::THREADS) {
        const int l = Div24<M, Z::LPB * M>::div(i), mm = i - l * M;
        finish(i, l, mm, s.a[l * Z::LSA + mm]);
      }
    }
    if (sm.peak_idx >= 0) {
      // one atomic per wave, spread over the period's slots by wave id
      for (int off = 32; off > 0; off >>= 1) pk = max(pk, (unsigned)__shfl_xor((int)pk, off));
      if ((threadIdx.x & 63) == 0 && pk) {
        const int slot = (blockIdx.x * ((Z::THREADS + 63) / 64) + (threadIdx.x >> 6)) % kPeakSlots;
        atomicMax(peaksp + (long)sm.peak_idx * kPeakWords + slot * kPeakStride, pk);
      }
    }
  } else {
    ZStage<M, ZV> stg;
    z_stage_load<M, ZV>(twg, rtg, m, a.s, line0, nl, stg);
    z_vol_load<M, ZV>(p, nl, pp);
    z_stage_put<M, ZV>(stg, a.s, line0, nl, s);
#pragma unroll
    for (int j = 0; j < CN::EPT; ++j) {
      const int l = ZElem<M, ZV>::line(j), mm = ZElem<M, ZV>::pos(j, l);
      if (l < nl) s.a[l * Z::LSA + mm] = pp[j];
    }
  }
  __syncthreads();
  z_r2c<M, ZV>(ll, t, lvalid, s, G(m.sp) + s.row[lvalid ? ll : 0], G(m.sp) + a.s.so + line0 + (lvalid ? ll : 0));
}

// ================================================================ launch dispatch
// Tile variants of the strided passes (Plan<L, TV>): instantiated for lengths >= 64,
// where the tile width has room to change; f receives the variant as a constant.
template <int N, class F>
inline void with_tile_variant(int tv, F f) {
  if constexpr (N >= 64) {
    if (tv == 1) return f(std::integral_constant<int, 1>{});
    if (tv == 2) return f(std::integral_constant<int, 2>{});
  }
  f(std::integral_constant<int, 0>{});
}

// z-pass variants (ZPlan<M, V>): layout (V & 3) among `layouts` and radix plan (V >> 2),
// the second radix plan only where it differs from the first (A != B).
template <int N, int NLAYOUT, class F>
inline void with_z_variant(int v, F f) {
  constexpr bool two = plan_a(N) * plan_a(N) != N;
  const int lay = (v & 3) < NLAYOUT ? (v & 3) : 0;
  const bool r1 = two && (v >> 2) == 1;
  if constexpr (two) {
    if (r1) {
      if (lay == 1) return f(std::integral_constant<int, 5>{});
      if constexpr (NLAYOUT > 2)
        if (lay == 2) return f(std::integral_constant<int, 6>{});
      return f(std::integral_constant<int, 4>{});
    }
  }
  if (lay == 1) return f(std::integral_constant<int, 1>{});
  if constexpr (NLAYOUT > 2)
    if (lay == 2) return f(std::integral_constant<int, 2>{});
  f(std::integral_constant<int, 0>{});
}

}  // namespace cst
