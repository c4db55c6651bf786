// Small-DFT codelets and compile-time twiddles for the Stockham line FFTs.
//
// Conventions: forward transform X[k] = sum_n x[n] exp(-2 pi i n k / N) (DIR = -1),
// inverse is the unnormalized conjugate kernel (DIR = +1), as numpy/scipy use.
// Codelets work on a register array float2 v[N] in place and return natural order.
// Composite sizes use one Cooley-Tukey split N = A * B with n = n1 + A n2,
// k = k2 + B k1:  B-point DFTs over n2, twiddle W_N^(n1 k2), A-point DFTs over n1.
// Twiddles are compile-time constants from an exact-octant-reduced Taylor series
// (double precision), so every multiply by 1, -1, +-i disappears at compile time.
#pragma once
#include <hip/hip_runtime.h>

namespace cst {

constexpr double kPi = 3.141592653589793238462643383279502884;

// sin/cos of an angle in [0, pi/4] by Taylor series (20 terms: far below double ulp).
constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int n = 1; n < 20; ++n) {
    term *= -x * x / double((2 * n) * (2 * n + 1));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int n = 1; n < 20; ++n) {
    term *= -x * x / double((2 * n - 1) * (2 * n));
    sum += term;
  }
  return sum;
}

// cos and sin of 2 pi num / den, reduced to the first octant with integer
// arithmetic so that multiples of pi/4 come out exact.
struct SinCos {
  double c, s;
};
constexpr SinCos sincos_turn(long num, long den) {
  long r = ((num % den) + den) % den;  // angle = 2 pi r / den in [0, 2 pi)
  // work in units of 1/(8 den) of a turn: a = 8 r, octant = a / den
  long a = 8 * r;
  long oct = a / den;         // 0..7
  long rem = a - oct * den;   // angle within octant, in units of turn/(8 den)
  // phi in [0, pi/4]: phi = 2 pi rem / (8 den)
  double phi = 2.0 * kPi * double(rem) / double(8 * den);
  double s0 = taylor_sin(phi), c0 = taylor_cos(phi);
  // complementary angle pi/4 - phi for odd octants
  double psi = 2.0 * kPi * double(den - rem) / double(8 * den);
  double s1 = taylor_sin(psi), c1 = taylor_cos(psi);
  if (rem == 0) { s0 = 0.0; c0 = 1.0; }
  if (rem == 0) { s1 = taylor_sin(kPi / 4); c1 = taylor_cos(kPi / 4); }
  switch (oct) {
    case 0: return {c0, s0};                // theta = phi
    case 1: return {s1, c1};                // theta = pi/2 - psi
    case 2: return {-s0, c0};               // theta = pi/2 + phi
    case 3: return {-c1, s1};               // theta = pi - psi
    case 4: return {-c0, -s0};              // theta = pi + phi
    case 5: return {-s1, -c1};              // theta = 3pi/2 - psi
    case 6: return {s0, -c0};               // theta = 3pi/2 + phi
    default: return {c1, -s1};              // theta = 2pi - psi
  }
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulconj(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by i * sign
template <int SIGN>
__device__ __forceinline__ float2 mul_i(float2 a) {
  return SIGN > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// v * W_den^(num) with W = exp(DIR * 2 pi i / den); special angles become exact ops.
template <int DIR, long NUM, long DEN>
__device__ __forceinline__ float2 twiddle_const(float2 v) {
  constexpr long r = ((NUM % DEN) + DEN) % DEN;
  if constexpr (r == 0) {
    return v;
  } else if constexpr (4 * r == DEN) {
    return mul_i<DIR>(v);
  } else if constexpr (2 * r == DEN) {
    return make_float2(-v.x, -v.y);
  } else if constexpr (4 * r == 3 * DEN) {
    return mul_i<-DIR>(v);
  } else {
    constexpr SinCos sc = sincos_turn(NUM, DEN);
    constexpr float c = float(sc.c);
    constexpr float s = float(DIR) * float(sc.s);
    return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
  }
}

// ---------------------------------------------------------------- codelets
template <int N, int DIR>
struct Dft;

template <int DIR>
struct Dft<1, DIR> {
  __device__ __forceinline__ static void run(float2 (&)[1]) {}
};

template <int DIR>
struct Dft<2, DIR> {
  __device__ __forceinline__ static void run(float2 (&v)[2]) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int DIR>
struct Dft<3, DIR> {
  __device__ __forceinline__ static void run(float2 (&v)[3]) {
    constexpr float h = 0.866025403784438646763723170752936183f;  // sqrt(3)/2
    float2 a = v[0], b = v[1], c = v[2];
    float2 t1 = cadd(b, c);
    float2 t2 = make_float2(a.x - 0.5f * t1.x, a.y - 0.5f * t1.y);
    float2 d = csub(b, c);
    float2 t3 = mul_i<DIR>(cscale(d, h));
    v[0] = cadd(a, t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
  }
};

template <int DIR>
struct Dft<4, DIR> {
  __device__ __forceinline__ static void run(float2 (&v)[4]) {
    float2 a = v[0], b = v[1], c = v[2], d = v[3];
    float2 s0 = cadd(a, c), d0 = csub(a, c);
    float2 s1 = cadd(b, d), d1 = mul_i<DIR>(csub(b, d));
    v[0] = cadd(s0, s1);
    v[2] = csub(s0, s1);
    v[1] = cadd(d0, d1);
    v[3] = csub(d0, d1);
  }
};

template <int DIR>
struct Dft<5, DIR> {
  __device__ __forceinline__ static void run(float2 (&v)[5]) {
    constexpr SinCos w1 = sincos_turn(1, 5);
    constexpr SinCos w2 = sincos_turn(2, 5);
    constexpr float c1 = float(w1.c), s1 = float(w1.s);
    constexpr float c2 = float(w2.c), s2 = float(w2.s);
    float2 a0 = v[0];
    float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    float2 r1 = make_float2(a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y);
    float2 r2 = make_float2(a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y);
    float2 q1 = mul_i<DIR>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
    float2 q2 = mul_i<DIR>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
    v[0] = cadd(a0, cadd(t1, t2));
    v[1] = cadd(r1, q1);
    v[4] = csub(r1, q1);
    v[2] = cadd(r2, q2);
    v[3] = csub(r2, q2);
  }
};

constexpr int first_factor(int n) {
  return (n % 4 == 0 && n != 4) ? 4 : (n % 2 == 0 ? 2 : (n % 3 == 0 ? 3 : 5));
}

template <int N, int DIR, int A, int B>
struct DftSplit {
  template <int N1, int K2>
  __device__ __forceinline__ static void tw_one(float2 (&y)[N]) {
    y[N1 + A * K2] = twiddle_const<DIR, long(N1) * K2, N>(y[N1 + A * K2]);
  }
  template <int N1, int... K2>
  __device__ __forceinline__ static void tw_row(float2 (&y)[N], std::integer_sequence<int, K2...>) {
    (tw_one<N1, K2>(y), ...);
  }
  template <int... N1>
  __device__ __forceinline__ static void tw_all(float2 (&y)[N], std::integer_sequence<int, N1...>) {
    (tw_row<N1>(y, std::make_integer_sequence<int, B>{}), ...);
  }
  __device__ __forceinline__ static void run(float2 (&v)[N]) {
    // step 1: B-point DFTs over n2 for every n1; result Y[n1][k2] kept at n1 + A k2
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) {
      float2 t[B];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) t[n2] = v[n1 + A * n2];
      Dft<B, DIR>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) v[n1 + A * k2] = t[k2];
    }
    tw_all(v, std::make_integer_sequence<int, A>{});
    // step 2: A-point DFTs over n1 for every k2; output X[k2 + B k1]
    float2 out[N];
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) {
      float2 t[A];
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) t[n1] = v[n1 + A * k2];
      Dft<A, DIR>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) out[k2 + B * k1] = t[k1];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = out[i];
  }
};

template <int N, int DIR>
struct Dft : DftSplit<N, DIR, first_factor(N), N / first_factor(N)> {};

}  // namespace cst
