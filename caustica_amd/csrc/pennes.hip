// Pennes bioheat + CEM43 dose: one fused explicit step per launch (k_pennes), or two
// (k_pennes2).
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
#include <mutex>
#include <string>
#include <type_traits>

#include "../../include/caustica_hip.h"
#include "common.h"
#include "thermal_state.h"

using namespace cst;

// Scratch of cst_thermal_run_noalloc (see the entry point): the second theta buffer of
// the ping-pong (kNoallocMaxVoxels floats) and the non-finite flag after it.
constexpr long kNoallocMaxVoxels = 1L << 24;  // 256^3
constexpr size_t kScratchBytes = (size_t)kNoallocMaxVoxels * 4 + 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static_assert(kNoallocMaxVoxels == CST_NOALLOC_MAX_VOXELS, "limit stated in the header");

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
  int xch2;  // x planes per block of k_pennes2
  // k, perfusion and dt/(rho C) of every voxel when all three maps are uniform (the
  // UNI kernels take these and never read the maps)
  float ku, pfu, cfu;
};

// Maps whose every element has the bits of element 0 leave *flag as it was; any other
// map sets it to 1 (a plain store: every writer stores the same value).
__global__ void __launch_bounds__(256) k_uniform_check(const unsigned* __restrict__ m, long n, unsigned* flag) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  const bool differs = i < n && m[i] != m[0];
  if (__syncthreads_or(differs) && threadIdx.x == 0) flag[0] = 1u;
}

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

// one explicit step of a voxel at (x, y, z), its neighbours' theta and k given: the update
// both kernels run
__device__ __forceinline__ float pennes_new_theta(const ThermalArgs& a, int x, int y, int z, float tc,
                                                  float kc, float tm, float tp, float kfm, float kfp,
                                                  float tym, float kym, float typ, float kyp, float tzm,
                                                  float kzm, float tzp, float kzp, float q, float pf,
                                                  float cf, float iw) {
  float work = 0.f;
  if (x > 0) work += kfm * iw * (tm - tc);
  else if (a.dirichlet & 1) work += 2.f * kc * iw * (a.theta_wall - tc);
  if (x + 1 < a.nx) work += kfp * iw * (tp - tc);
  else if (a.dirichlet & 1) work += 2.f * kc * iw * (a.theta_wall - tc);
  if (y > 0) work += hmean(kym, kc) * iw * (tym - tc);
  else if (a.dirichlet & 2) work += 2.f * kc * iw * (a.theta_wall - tc);
  if (y + 1 < a.ny) work += hmean(kc, kyp) * iw * (typ - tc);
  else if (a.dirichlet & 2) work += 2.f * kc * iw * (a.theta_wall - tc);
  if (z > 0) work += hmean(kzm, kc) * iw * (tzm - tc);
  else if (a.dirichlet & 4) work += 2.f * kc * iw * (a.theta_wall - tc);
  if (z + 1 < a.nz) work += hmean(kc, kzp) * iw * (tzp - tc);
  else if (a.dirichlet & 4) work += 2.f * kc * iw * (a.theta_wall - tc);
  work += q;
  work -= pf * tc;
  return tc + work * cf;
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
// UNI: k, perfusion and dt/(rho C) are the scalars a.ku, a.pfu, a.cfu (all three maps
// uniform); the arithmetic is the map kernel's, on the same values, so the results are
// the same bits.
// theta_max is stored only where it changes (a masked store leaves lines no lane writes
// untouched in memory: most of a cooling volume).
template <bool HAS_Q, bool DOSE, bool FULL, bool UNI>
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
  auto kk = [&](int x, long c) { return UNI ? a.ku : a.k[(long)x * plane + c]; };
  // loads for plane x: its halo and coefficients, and the centre of plane x+1
  auto prefetch = [&](int x, PennesPre& p) {
    const int xc = min(x, xlast), xp = min(x + 1, xlast);
    const long i = (long)xc * plane + col;
    p.tp = th(xp, col);
    p.kp = kk(xp, col);
    p.ht = th(xc, colh);
    p.hk = kk(xc, colh);
    p.q = HAS_Q ? a.q[i] : 0.f;
    p.pf = UNI ? a.pfu : a.perf[i];
    p.cf = UNI ? a.cfu : a.coeff[i];
    p.tx = a.tmax[i];
    p.dz = DOSE ? a.dose[i] : 0.0;
  };
  // k of the plane's y/z neighbours (from the LDS tile unless uniform)
  auto sK = [&](int b, int yy, int zz) { return UNI ? a.ku : sk[b][yy][zz]; };
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
    if (!UNI) sk[b][ty + 1][tz + 1] = kc;
    if (hv) {
      sth[b][hly][hlz] = c.ht;
      if (!UNI) sk[b][hly][hlz] = c.hk;
    }
    __syncthreads();
    const float tp = c.tp, kp = c.kp;
    const float kfp = hmean(kc, kp);  // x face above plane x
    if (valid) {
      const long i = (long)x * plane + col;
      const float tn = pennes_new_theta(
          a, x, y, z, tc, kc, tm, tp, kfm, kfp, sth[b][ty][tz + 1], sK(b, ty, tz + 1), sth[b][ty + 2][tz + 1],
          sK(b, ty + 2, tz + 1), sth[b][ty + 1][tz], sK(b, ty + 1, tz), sth[b][ty + 1][tz + 2],
          sK(b, ty + 1, tz + 2), HAS_Q ? c.q : 0.f, c.pf, c.cf, iw);
      a.th_out[i] = tn;
      const float mx = fmaxf(c.tx, tn);
      if (__float_as_uint(mx) != __float_as_uint(c.tx)) a.tmax[i] = mx;
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

// ---------------------------------------------------------------- two steps per sweep
// theta n -> n+2 in one pass over memory: per voxel and two steps, theta is read and
// written once, k, Q, perfusion and dt/(rho C) read once, theta_max and the dose read
// and written once: 48 P bytes per two steps instead of 96. Each block computes step n+1
// on its TZ x TY2 tile plus a one-cell ring (its neighbours' tiles recompute that ring
// too), from theta^n and k on the tile plus two rings, and step n+2 on the tile one
// plane behind. Both steps run k_pennes's update (pennes_new_theta) and take theta_max
// and the dose in the same order: two steps come out as two launches of k_pennes give them.
constexpr int TY2 = 8;                             // tile rows (y): 64 x 8, eight waves
constexpr int RW1 = TZ + 2, RW2 = TZ + 4;          // row lengths with one and two rings
constexpr int NR1 = 2 * RW1 + 2 * TY2;             // cells of ring 1 (148)
constexpr int NR2 = 2 * RW2 + 2 * (TY2 + 2);       // cells of ring 2 (156)
static_assert(NR1 <= TZ * TY2 && NR2 <= TZ * TY2, "one ring-1 and one ring-2 cell per thread at most");

// (dy, dz) of cell i of ring r (1 or 2) relative to the tile's first cell: the rows above
// and below first, then the two ends of each row in between
__device__ __forceinline__ void ring_cell(int r, int i, int& dy, int& dz) {
  const int w = TZ + 2 * r;
  if (i < 2 * w) {
    const int side = i / w;
    dz = i - side * w - r;
    dy = side ? TY2 - 1 + r : -r;
  } else {
    const int j = i - 2 * w;
    dy = (j >> 1) - (r - 1);
    dz = (j & 1) ? TZ - 1 + r : -r;
  }
}

// x planes per block (a.xch2): the longest of 64, 32, 16, 8 that still gives the grid as
// many blocks per CU as a CU holds at once: two of the map kernels (84-93 VGPRs, five
// waves per SIMD), four of the uniform-map ones (48-57 VGPRs, eight). At 256^3: 64 planes
// (512 blocks) and 32 (1024 blocks), the two recomputed chunk-border planes 3 % and 6 %
// of the work; smaller grids take shorter chunks rather than leave CUs idle. The 64 x 8
// tile keeps the recomputed ring at 29 % of the tile (55 % for 64 x 4).

// Loads of one plane step, issued one iteration before they are consumed: theta, k of
// slots 0 and 1 at xa + 1, their maps at xa, theta, k of slot 2 at xa, theta_max and dose
// of the tile cell at xa - 1 (for step n+2 there). Two sets alternate (the march is
// unrolled twice), so consuming one never waits on the loads just issued into the other.
struct Pennes2Pre {
  float t[2], k[2], q[2], pf[2], cf[2];
  float t2, k2, tx;
  double dz;
};

// UNI and the theta_max store as in k_pennes.
template <bool HAS_Q, bool DOSE, bool UNI>
__global__ void __launch_bounds__(TZ* TY2) k_pennes2(ThermalArgs a) {
  __shared__ float sT[TY2 + 4][RW2];      // theta^n of plane xa on the tile + 2 rings
  __shared__ float sKm[2][TY2 + 4][RW2];  // k of planes xa (b) and xa - 1 (1 - b)
  auto sK = [&](int b, int yy, int zz) { return UNI ? a.ku : sKm[b][yy][zz]; };
  __shared__ float sU[2][TY2 + 2][RW1];   // theta^n+1 of planes xa (b) and xa - 1 (1 - b)
  const int tid = threadIdx.x, tz = tid % TZ, ty = tid / TZ;
  const int y0 = blockIdx.y * TY2, z0 = blockIdx.x * TZ;
  const int x0 = blockIdx.z * a.xch2, x1 = min(x0 + a.xch2, a.nx);
  const long plane = (long)a.ny * a.nz;
  const float iw = a.inv_dx2;
  // the thread's cells: slot 0 the tile cell, slot 1 a ring-1 cell (threads 0 .. NR1 - 1),
  // slot 2 a ring-2 cell (the second half of the block: its loads and LDS writes go to
  // other waves than the ring-1 updates)
  constexpr int H2 = TZ * TY2 / 2;
  static_assert(H2 >= NR1 && H2 >= NR2, "ring cells of the two halves do not share a thread");
  int cy[3], cz[3];
  cy[0] = ty;
  cz[0] = tz;
  const bool has1 = tid < NR1, has2 = tid >= H2 && tid < H2 + NR2;
  ring_cell(1, has1 ? tid : 0, cy[1], cz[1]);
  ring_cell(2, has2 ? tid - H2 : 0, cy[2], cz[2]);
  bool in[3];
  long col[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int y = y0 + cy[c], z = z0 + cz[c];
    in[c] = y >= 0 && y < a.ny && z >= 0 && z < a.nz;
    col[c] = (long)min(max(y, 0), a.ny - 1) * a.nz + min(max(z, 0), a.nz - 1);
  }
  in[1] = in[1] && has1;
  in[2] = in[2] && has2;
  const int xlast = a.nx - 1;
  auto F = [&](const float* m, int x, long c) { return m[(long)min(max(x, 0), xlast) * plane + c]; };
  auto K = [&](int x, long c) { return UNI ? a.ku : F(a.k, x, c); };
  auto prefetch = [&](int xa, Pennes2Pre& p) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      p.t[c] = F(a.th_in, xa + 1, col[c]);
      p.k[c] = K(xa + 1, col[c]);
      p.q[c] = HAS_Q ? F(a.q, xa, col[c]) : 0.f;
      p.pf[c] = UNI ? a.pfu : F(a.perf, xa, col[c]);
      p.cf[c] = UNI ? a.cfu : F(a.coeff, xa, col[c]);
    }
    p.t2 = F(a.th_in, xa, col[2]);
    p.k2 = K(xa, col[2]);
    p.tx = F(a.tmax, xa - 1, col[0]);
    p.dz = DOSE ? a.dose[(long)min(max(xa - 1, 0), xlast) * plane + col[0]] : 0.0;
  };
  // carried step-n state of slots 0 and 1: theta at xa - 1 and xa, k at xa, x face below xa
  float tm[2], tc[2], kc[2], kfm[2];
  const int xs = x0 - 1;  // first plane of step n+1
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    tm[c] = F(a.th_in, xs - 1, col[c]);
    tc[c] = F(a.th_in, xs, col[c]);
    kc[c] = K(xs, col[c]);
    kfm[c] = hmean(K(xs - 1, col[c]), kc[c]);
  }
  Pennes2Pre P0, P1;
  prefetch(xs, P0);
  // step n+2 state of the tile cell at xb = xa - 1: theta^n+1 at xb - 1 and xb, theta^n,
  // k, x faces and maps at xb
  float u_m = 0.f, u_c = 0.f, tn_b = 0.f, kc_b = 0.f, fm_b = 0.f, fp_b = 0.f;
  float q_b = 0.f, pf_b = 0.f, cf_b = 0.f;
  bool bad = false;
  auto iter = [&](int xa, int b, const Pennes2Pre& cur, Pennes2Pre& nxt) {
    const int xb = xa - 1;
    prefetch(xa + 1, nxt);
    // theta^n and k of plane xa on the tile and both rings
    sT[cy[0] + 2][cz[0] + 2] = tc[0];
    if (!UNI) sKm[b][cy[0] + 2][cz[0] + 2] = kc[0];
    if (has1) {
      sT[cy[1] + 2][cz[1] + 2] = tc[1];
      if (!UNI) sKm[b][cy[1] + 2][cz[1] + 2] = kc[1];
    }
    if (has2) {
      sT[cy[2] + 2][cz[2] + 2] = cur.t2;
      if (!UNI) sKm[b][cy[2] + 2][cz[2] + 2] = cur.k2;
    }
    __syncthreads();
    // step n+1 at plane xa for slots 0 and 1
    const bool xa_in = xa >= 0 && xa < a.nx;
    const float kfm0 = kfm[0];
    float u_new = 0.f, kfp0 = 0.f;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const float kfp = hmean(kc[c], cur.k[c]);
      if (c == 0) kfp0 = kfp;
      if (xa_in && in[c]) {
        const int Y = cy[c] + 2, Z = cz[c] + 2;
        const float un = pennes_new_theta(
            a, xa, y0 + cy[c], z0 + cz[c], tc[c], kc[c], tm[c], cur.t[c], kfm[c], kfp, sT[Y - 1][Z],
            sK(b, Y - 1, Z), sT[Y + 1][Z], sK(b, Y + 1, Z), sT[Y][Z - 1], sK(b, Y, Z - 1), sT[Y][Z + 1],
            sK(b, Y, Z + 1), cur.q[c], cur.pf[c], cur.cf[c], iw);
        sU[b][cy[c] + 1][cz[c] + 1] = un;
        if (c == 0) u_new = un;
      }
      kfm[c] = kfp;
    }
    // no barrier: step n+2 reads only what the previous plane step wrote (sU[1 - b],
    // sK[1 - b]) and its own cell's u_new, so the waves without a ring cell go on while
    // the others finish theirs
    // step n+2 at plane xb for the tile cell (theta^n+1 of plane xb in sU[1 - b], of
    // planes xb -+ 1 in u_m and u_new)
    if (xb >= x0 && xb < x1 && in[0]) {
      const int Y = cy[0] + 1, Z = cz[0] + 1, YK = cy[0] + 2, ZK = cz[0] + 2;
      const float t2n = pennes_new_theta(
          a, xb, y0 + cy[0], z0 + cz[0], u_c, kc_b, u_m, u_new, fm_b, fp_b, sU[1 - b][Y - 1][Z],
          sK(1 - b, YK - 1, ZK), sU[1 - b][Y + 1][Z], sK(1 - b, YK + 1, ZK), sU[1 - b][Y][Z - 1],
          sK(1 - b, YK, ZK - 1), sU[1 - b][Y][Z + 1], sK(1 - b, YK, ZK + 1), q_b, pf_b, cf_b, iw);
      const long i = (long)xb * plane + col[0];
      a.th_out[i] = t2n;
      const float mx = fmaxf(fmaxf(cur.tx, u_c), t2n);
      if (__float_as_uint(mx) != __float_as_uint(cur.tx)) a.tmax[i] = mx;
      bad |= !isfinite(u_c) || !isfinite(t2n);
      if (DOSE) {
        const float r0 = cem43_rate(tn_b + a.t_ref);
        const float r1 = cem43_rate(u_c + a.t_ref);
        const float r2 = cem43_rate(t2n + a.t_ref);
        const double d1 = cur.dz + ((double)r1 + (double)r0) * a.half_dt_min;
        const double d2 = d1 + ((double)r2 + (double)r1) * a.half_dt_min;
        a.dose[i] = d2;
        bad |= !isfinite(d1) || !isfinite(d2);
      }
    }
    __syncthreads();  // sT and sK[1 - b] are rewritten next
    // the tile cell's plane-xa values become step n+2's plane xb next
    u_m = u_c;
    u_c = u_new;
    tn_b = tc[0];
    kc_b = kc[0];
    fm_b = kfm0;
    fp_b = kfp0;
    q_b = cur.q[0];
    pf_b = cur.pf[0];
    cf_b = cur.cf[0];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      tm[c] = tc[c];
      tc[c] = cur.t[c];
      kc[c] = cur.k[c];
    }
  };
  for (int xa = xs; xa <= x1; xa += 2) {
    iter(xa, 0, P0, P1);
    if (xa + 1 <= x1) iter(xa + 1, 1, P1, P0);
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

// f(std::integral_constant<bool, a>, ..., <bool, c>): three run-time flags as template
// arguments
template <typename F>
void with_flags(bool a, bool b, bool c, F f) {
  auto pick = [](bool v, auto g) {
    if (v) g(std::true_type{});
    else g(std::false_type{});
  };
  pick(a, [&](auto A) { pick(b, [&](auto B) { pick(c, [&](auto C) { f(A, B, C); }); }); });
}

// n explicit sub-steps ping-ponging theta between buf[cur] and buf[1 - cur] (cur is
// updated): pairs of sub-steps as one two-step sweep (k_pennes2), a remainder as one-step
// launches (k_pennes). uni: the UNI kernels (a.ku, a.pfu, a.cfu hold the uniform maps'
// value). The launches only: nothing is allocated or synchronized.
int launch_substeps(ThermalArgs a, float* const buf[2], int& cur, int n_substeps, int device,
                    hipStream_t stream, Profiler* prof, int xch2 = 0, bool uni = false) {
  dim3 grid((a.nz + TZ - 1) / TZ, (a.ny + TY - 1) / TY, (a.nx + XCH - 1) / XCH);
  int s = 0;
  // pairs of sub-steps: two steps per sweep over memory (k_pennes2)
  int ncu = 0;
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
  const long yz2 = (long)((a.nz + TZ - 1) / TZ) * ((a.ny + TY2 - 1) / TY2);
  a.xch2 = 64;
  const long resident = (uni ? 4L : 2L) * ncu;
  while (a.xch2 > 8 && yz2 * ((a.nx + a.xch2 - 1) / a.xch2) < resident) a.xch2 /= 2;
  if (xch2 > 0) a.xch2 = xch2;  // the caller's chunk length (cst_thermal_set_chunk)
  const dim3 grid2((a.nz + TZ - 1) / TZ, (a.ny + TY2 - 1) / TY2, (a.nx + a.xch2 - 1) / a.xch2);
  auto timed = [&](auto f) {
    if (prof) prof->run(0, stream, f);
    else f();
  };
  const bool hq = a.q != nullptr, dd = a.do_dose != 0, full = a.nx % XCH == 0;
  for (; s + 2 <= n_substeps; s += 2) {
    a.th_in = buf[cur];
    a.th_out = buf[1 - cur];
    timed([&] {
      with_flags(hq, dd, uni, [&](auto HQ, auto DD, auto U) {
        hipLaunchKernelGGL((k_pennes2<decltype(HQ)::value, decltype(DD)::value, decltype(U)::value>), grid2,
                           dim3(TZ * TY2), 0, stream, a);
      });
      return true;
    });
    HIPCHK(hipGetLastError());
    cur = 1 - cur;
  }
  for (; s < n_substeps; ++s) {
    a.th_in = buf[cur];
    a.th_out = buf[1 - cur];
    timed([&] {
      with_flags(hq, dd, uni, [&](auto HQ, auto DD, auto U) {
        if (full)
          hipLaunchKernelGGL((k_pennes<decltype(HQ)::value, decltype(DD)::value, true, decltype(U)::value>), grid,
                             dim3(TZ * TY), 0, stream, a);
        else
          hipLaunchKernelGGL((k_pennes<decltype(HQ)::value, decltype(DD)::value, false, decltype(U)::value>),
                             grid, dim3(TZ * TY), 0, stream, a);
      });
      return true;
    });
    HIPCHK(hipGetLastError());
    cur = 1 - cur;
  }
  return CST_OK;
}

// ------------------------------------------------------------ no-allocation entry point
// Scratch of cst_thermal_run_noalloc: the second theta buffer of the ping-pong and the
// non-finite flag, in a file-scope device array (every GPU the library runs on holds its
// own copy until the process ends). Host code never uses the array's name as an address:
// each device's address comes from hipGetSymbolAddress on the entry point's first call on
// that device, and only that address reaches kernels.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_for(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return fail(CST_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CST_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return CST_OK;
}

__global__ void k_flag_clear(unsigned* f) {
  if (threadIdx.x == 0) f[0] = 0u;
}

__global__ void __launch_bounds__(256) k_copy_f32(const float* __restrict__ src, float* __restrict__ dst,
                                                  long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
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
  // bad[0]: the non-finite flag; bad[1..3]: k, perfusion, dt/(rho C) not uniform
  HIPCHK(hipMemsetAsync(t->bad, 0, 16, st));
  const float* maps[3] = {t->k, t->perf, t->coeff};
  for (int m = 0; m < 3; ++m)
    hipLaunchKernelGGL(k_uniform_check, dim3(blocks), dim3(thr), 0, st, reinterpret_cast<const unsigned*>(maps[m]),
                       P, t->bad + 1 + m);
  HIPCHK(hipGetLastError());
  unsigned differs[3] = {1u, 1u, 1u};
  HIPCHK(hipMemcpyAsync(differs, t->bad + 1, sizeof differs, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  t->uniform = !differs[0] && !differs[1] && !differs[2];
  t->ku = d->k[0];
  t->pfu = d->perf[0];
  t->cfu = d->coeff[0];
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
  const bool uni = t->uniform && t->allow_uniform;
  a.ku = t->ku;
  a.pfu = t->pfu;
  a.cfu = t->cfu;
  return launch_substeps(a, t->theta, t->cur, n_substeps, t->device, t->stream, &t->prof, t->xch2, uni);
}

int cst_thermal_set_uniform_maps(cst_thermal* t, int allow, int* uniform) {
  if (!t) return fail(CST_ERR_ARG, "null argument");
  t->allow_uniform = allow != 0;
  if (uniform) *uniform = t->ready && t->uniform ? 1 : 0;
  return CST_OK;
}

int cst_thermal_set_chunk(cst_thermal* t, int planes) {
  if (!t) return fail(CST_ERR_ARG, "null argument");
  if (planes != 0 && planes != 8 && planes != 16 && planes != 32 && planes != 64)
    return fail(CST_ERR_ARG, "x planes per block must be 0 (chosen per grid), 8, 16, 32 or 64");
  t->xch2 = planes;
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

int cst_thermal_read_source(cst_thermal* t, float* q) {
  if (!t || !q) return fail(CST_ERR_ARG, "null argument");
  if (!t->ready) return fail(CST_ERR_STATE, "thermal session not set up");
  if (!t->have_q) return fail(CST_ERR_STATE, "thermal session has no heat source");
  const long P = (long)t->nx * t->ny * t->nz;
  HIPCHK(hipSetDevice(t->device));
  HIPCHK(hipMemcpyAsync(q, t->q, P * 4, hipMemcpyDeviceToHost, t->stream));
  HIPCHK(hipStreamSynchronize(t->stream));
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

int cst_thermal_run_noalloc(const cst_thermal_dev_desc* d, int n_substeps, void* stream, int* nonfinite) {
  if (!d) return fail(CST_ERR_ARG, "null argument");
  if (d->nx < 1 || d->ny < 1 || d->nz < 1) return fail(CST_ERR_ARG, "invalid thermal shape");
  const long P = (long)d->nx * d->ny * d->nz;
  if (P > kNoallocMaxVoxels)
    return fail(CST_ERR_ARG, "grid of " + std::to_string(P) + " voxels exceeds the no-allocation limit of " +
                                 std::to_string(kNoallocMaxVoxels) + " voxels (use the thermal session)");
  if (!d->theta || !d->theta_max || !d->k || !d->perf || !d->coeff || (d->dose_on && !d->dose))
    return fail(CST_ERR_ARG, "missing device arrays");
  if (!(d->dx > 0)) return fail(CST_ERR_ARG, "dx must be > 0");
  if (n_substeps < 0) return fail(CST_ERR_ARG, "negative step count");
  int device = 0;
  HIPCHK(hipGetDevice(&device));
  void* base = nullptr;
  if (int r = scratch_for(device, &base)) return r;
  float* other = static_cast<float*>(base);
  unsigned* flag = reinterpret_cast<unsigned*>(static_cast<unsigned char*>(base) + (size_t)kNoallocMaxVoxels * 4);
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_flag_clear, dim3(1), dim3(64), 0, st, flag);
  HIPCHK(hipGetLastError());
  ThermalArgs a{};
  a.tmax = d->theta_max;
  a.dose = d->dose;
  a.k = d->k;
  a.perf = d->perf;
  a.coeff = d->coeff;
  a.q = d->q;
  a.bad = flag;
  a.nx = d->nx;
  a.ny = d->ny;
  a.nz = d->nz;
  a.inv_dx2 = (float)(1.0 / (d->dx * d->dx));
  a.dirichlet = d->boundary;
  a.theta_wall = d->theta_wall;
  a.t_ref = d->t_ref;
  a.half_dt_min = d->half_dt_min;
  a.do_dose = d->dose_on;
  float* const buf[2] = {d->theta, other};
  int cur = 0;
  if (int r = launch_substeps(a, buf, cur, n_substeps, device, st, nullptr)) return r;
  if (cur == 1) {
    hipLaunchKernelGGL(k_copy_f32, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, other, d->theta, P);
    HIPCHK(hipGetLastError());
  }
  if (nonfinite) {
    unsigned v = 0;
    HIPCHK(hipMemcpyAsync(&v, flag, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    *nonfinite = v ? 1 : 0;
  }
  return CST_OK;
}

}  // extern "C"
