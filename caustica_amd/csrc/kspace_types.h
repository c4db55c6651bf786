// Device-side descriptors shared by the k-space pass kernels and the host session.
//
// Layout in HBM for one solve ("member") on a padded grid nx x ny x nz (z contiguous):
//   real volumes   [nx][ny][nz]      float     p, u0, u1, u2, dt/rho, rho c^2 dt, absorb, 2 beta dt
//   half spectra   [nx][ps]: [ny][ms] float2 (ms = nz/2+1 rounded up to 16 bins,
//                                              ps = ny ms + 16)    sp (p-hat, then the divergence spectrum),
//                                              sg0..sg2 (gradient spectra, then u-hat)
//   record box     [rx][ry][rz]      float2    Kahan sum + compensation per harmonic; float p_max
// Several members of one shape share every launch (blockIdx.z indexes the member list).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace cst {

constexpr int kMaxBatch = 16;     // members per launch
constexpr int kMaxHarmonics = 3;  // harmonics recorded per member
// Per-period convergence peaks are spread over kPeakSlots words, kPeakStride uints
// apart (64 bytes), so that blocks do not all hit one address with their atomicMax.
constexpr int kPeakSlots = 64;
constexpr int kPeakStride = 16;
constexpr int kPeakWords = kPeakSlots * kPeakStride;  // per period

// Arrays of a member of the opt-in staggered scheme (session scheme CST_SCHEME_STAGGERED):
// split density state rho_a and the maps and tables of its update. Null for members of
// the reference scheme.
struct StgDev {
  float* rho[3];           // split density rho_a (real volumes); sum_a rho_a = rho
  const float* dtrho[3];   // dt / rho0 at x_a + dx / 2 (velocity update of axis a)
  const float* dtrho0;     // dt rho0
  const float* c2;         // c^2
  const float* c2tau;      // c^2 tau / dt (absorbing members; tau < 0)
  const float* c2eta;      // c^2 eta (absorbing members)
  const float* nl2dt;      // 2 m dt, m = 1 where beta > 0 (null: linear)
  const float* c2bh;       // c^2 m (beta - 1) / rho0 (nonlinear members)
  const float2* kfac;      // k_a exp(+i k_a dx / 2): x[nx], y[ny], z[nz/2 + 1]
  const float* pml;        // 4 tables (sz[nz] sx[nx] sy[ny], pml_stride floats apart) of
                           // exp(-sigma dt / 2): 0 at grid points on every axis, 1 + a at
                           // staggered points x_a + dx / 2 along axis a, ones elsewhere
  int pml_stride;          // nz + nx + ny rounded up to a multiple of 64
  float e1, e2;            // exponents on |k|^2: (y - 2) / 2 and (y - 1) / 2
  int absorbing;           // fractional-Laplacian absorption terms on
  int src_axes;            // bit a set: the mass source adds to rho_a
};

// One member's descriptor, in a device table the kernels index by member slot. Entries
// are whole 256-byte units, so that writing one member's entry (an admission while
// other members' launches are in flight) never shares a cache line with another's.
struct alignas(256) MemberDev {
  float* p;
  float2* sp;
  float2* sg[3];
  // velocity state: the z-spectrum (unnormalized R2C along z, same row layout as the
  // other spectra) of each component; the real u is C2R(su) / nz, formed in the z pass
  float2* su[3];
  const float* dto_rho;
  const float* rhoc2dt;
  const float* absorb;
  const float* beta2dt;  // null: linear update
  // k tables: kx_full[nx] kx_d[nx] ky_full[ny] ky_d[ny] kz_full[nzh] kz_d[nzh]
  // (_d = derivative wavenumber with the even-axis Nyquist bin zeroed)
  const float* ktab;
  const float* sponge;   // sz[nz] sx[nx] sy[ny] (sz first: float2-aligned)
  const int* src_ptr;    // [nx*ny + 1] CSR over (x, y) lines
  const int* src_z;      // z index of each source voxel, grouped by line
  const float2* src_ab;  // (scale cos(phase), scale sin(phase)) per source voxel
  float2* acc[kMaxHarmonics];
  float2* comp[kMaxHarmonics];
  float* pmax;
  unsigned* peaks;       // [period][kPeakWords]: max |p| over the convergence box (float bits)
  int conv[6];           // x0, x1, y0, y1, z0, z1 (half-open)
  int rec[6];
  int nh;
  float half_cdt;        // c_ref * dt / 2
  float inv_n;           // 1 / (nx ny nz): the inverse-transform normalization
  float inv_nz;          // 1 / nz: the z round trip of the velocity state
  // Step clock, in device memory so that a time step's launches take the same
  // arguments at every step (a captured step replays as it is): clock[0] counts the
  // steps begun (the first pass of a step, k_xgrad, advances it), clock[1] is the
  // first recorded step. The step scalars are derived from it on the device
  // (step_scalars): drive amplitude from the ramp table, drive and harmonic phases from
  // the spp-entry (cos, sin) table of 2 pi k / spp built in float64 on the host. The
  // thread of k_xgrad that advances the clock stores them at stepv for the step's
  // later passes: one load there, instead of a chain of dependent loads and 64-bit
  // divisions at the start of every block.
  long long* clock;
  struct StepMember* stepv;
  const float* amp_tab;  // float32(amplitude * ramp envelope(n dt)) for n < ramp_n
  const float2* trig;    // (cos, sin)(2 pi k / spp), k < spp
  float amp_full;        // float32(amplitude) once the ramp is complete
  int ramp_n;
  int spp;
  int max_periods;
  int harm[kMaxHarmonics];
  StgDev stg;
};

// Global-address-space view of a device pointer read from a table. Loads and stores
// through it compile to global_* instructions; through a generic pointer they become
// flat_* instructions, which also count against the LDS wait counter, so every wait for
// an LDS result would first drain the HBM loads in flight.
// (The host pass only checks kernel bodies; it sees a plain pointer.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using gptr = __attribute__((address_space(1))) T*;
#else
template <class T>
using gptr = T*;
#endif
template <class T>
__device__ __forceinline__ gptr<T> G(T* p) {
  return (gptr<T>)p;
}
// value of p[i] (a conditional expression cannot mix the qualified lvalue with a float2)
template <class T>
__device__ __forceinline__ std::remove_const_t<T> ld(gptr<T> p, long i) {
  return p[i];
}

struct StepMember {
  float amp;      // float32(amplitude * ramp envelope at t = n dt)
  float sin_t;    // sin(2 pi (n mod spp) / spp)
  float cos_t;
  int record;     // accumulate phasors / p_max this step
  float2 eh[kMaxHarmonics];  // exp(+i h 2 pi (n mod spp) / spp)
  int peak_idx;   // period slot for the convergence peak, -1: none
};

// Step scalars of step n (engine.py:478-482 source drive, :535-550 record window,
// :490-495 period peak), record_from the member's first recorded step.
__device__ __forceinline__ StepMember step_scalars(const MemberDev& m, long long n, long long record_from) {
  StepMember sm;
  const int spp = m.spp;
  const int ph = (int)(n % spp);
  const gptr<const float2> trig = G(m.trig);
  const float2 cs = trig[ph];
  sm.cos_t = cs.x;
  sm.sin_t = cs.y;
  sm.amp = n < m.ramp_n ? G(m.amp_tab)[n] : m.amp_full;
  sm.record = n >= record_from ? 1 : 0;
#pragma unroll
  for (int h = 0; h < kMaxHarmonics; ++h) sm.eh[h] = trig[(int)(((long long)m.harm[h] * ph) % spp)];
  const long long pidx = n / spp;
  sm.peak_idx = (!sm.record && pidx < m.max_periods) ? (int)pidx : -1;
  return sm;
}

// Step scalars of the step in progress, as the step's k_xgrad stored them.
__device__ __forceinline__ StepMember member_step(const MemberDev& m) {
  return ld(G((const StepMember*)m.stepv), 0);
}

// Half-spectrum layout (nzh = nz/2 + 1 bins per (x, y)): bins k < nz/2 of (x, y) sit in
// rows at x ps + y ms + k; the Nyquist bin k = nz/2 of every (x, y) sits in a side plane
// at so + x ny + y. Rows thus hold a multiple of 16 bins (128 B) and every column tile
// of the strided passes is a full, aligned cache line; the side plane is one more small
// column for them.
struct Shape {
  int nx, ny, nz, nzh;
  int ms;   // row stride in bins: nz/2 rounded up to a multiple of 16
  long ps;  // x-plane stride in bins: ny * ms + 16 (breaks power-of-two strides)
  long so;  // offset of the Nyquist side plane: nx * ps
  long hs;  // bins per spectrum array: so + nx * ny rounded up to a multiple of 16
};

struct StepArgs {
  const MemberDev* mem;
  Shape s;
  int nmem;
  int ids[kMaxBatch];
};

}  // namespace cst
