// Host session of the CW k-space solver: device memory per member, step clocks,
// batched launch sequence (direct or replayed from captured HIP graphs), results,
// checkpoint state, kernel timing. C ABI in include/caustica_hip.h.
//
// Streams: every step of a session runs on its compute stream, in order; member
// uploads run on an upload stream (the host waits for them before cst_member_add
// returns, so nothing on the compute stream has to wait); read-backs (period peaks,
// finalized phasors) run on an io stream once the host has waited for the member's last
// queued step (an event recorded after each cst_advance), so that reading a finished
// member back overlaps the next steps of the others.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/caustica_hip.h"
#include "common.h"
#include "dl_abi.h"
#include "kspace_launch.h"
#include "kspace_types.h"
#include "thermal_state.h"

using namespace cst;

namespace {

struct Member {
  bool used = false;
  int64_t n = 0;
  int64_t record_from = INT64_MAX;
  double dt = 0, f0 = 0, amplitude = 0, ramp_periods = 3;
  int spp = 0, nh = 1, harmonics[kMaxHarmonics] = {1, 0, 0};
  int max_periods = 0;
  int64_t rec_count = 0;
  bool finalized[kMaxHarmonics] = {false, false, false};
  double final_scale[kMaxHarmonics] = {0, 0, 0};
  bool primed = true;
  void* block = nullptr;
  size_t block_bytes = 0;
  hipEvent_t done = nullptr;  // recorded after the member's last queued step
  MemberDev dev{};
};

// Sets a member's device step clock (which bit 0: steps begun, bit 1: first recorded
// step). The values travel as kernel arguments: no host buffer has to outlive the call.
__global__ void k_clock_set(long long* clock, long long n, long long record_from, int which) {
  if (threadIdx.x == 0) {
    if (which & 1) clock[0] = n;
    if (which & 2) clock[1] = record_from;
  }
}

__global__ void k_finalize(const float2* __restrict__ acc, float2* __restrict__ comp, int64_t n,
                           double scale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const float2 s = acc[i], c = comp[i];
    comp[i] = make_float2((float)(((double)s.x - (double)c.x) * scale),
                          (float)(((double)s.y - (double)c.y) * scale));
  }
}

struct HeatJoin {
  const float2* p[kMaxHarmonics];
  double f[kMaxHarmonics];
  int nt;
  const float* w;
  const float* perf;
  float qoff;
  double scale;
  int64_t n;
  float* q;
};

// q = float32(scale * w * sum_j f_j |P_j|^2) - perf * qoff (the subtraction in float32)
__global__ void __launch_bounds__(256) k_heat_join(HeatJoin j) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= j.n) return;
  double acc = 0.0;
  for (int t = 0; t < j.nt; ++t) {
    const float2 v = j.p[t][i];
    acc += j.f[t] * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
  }
  const float q = (float)(j.scale * (double)j.w[i] * acc);
  j.q[i] = __fsub_rn(q, __fmul_rn(j.perf[i], j.qoff));
}

}  // namespace

// A captured launch sequence of `steps` time steps for one member list.
struct GraphEntry {
  int nmem = 0;
  int ids[cst::kMaxBatch] = {};
  int v[4] = {0, 0, 0, 0};  // the session's pass variants at capture
  bool nonlinear = false;
  bool absorbing = false;  // staggered sessions: the absorption passes are in the sequence
  int steps = 0;
  hipGraphExec_t exec = nullptr;
  uint64_t last_use = 0;
};

struct cst_session {
  int device = 0;
  int scheme = CST_SCHEME_REFERENCE;
  Shape s{};
  int max_members = 0;
  hipStream_t stream = nullptr;
  hipStream_t up = nullptr;  // member uploads
  hipStream_t io = nullptr;  // read-backs
  bool graphs = true;
  std::vector<GraphEntry> gcache;
  uint64_t gclock = 0;
  int64_t graph_launches = 0, graph_captures = 0;
  std::vector<Member> members;
  MemberDev* d_table = nullptr;
  float2* d_tables = nullptr;  // twiddles: x (nx), y (ny), z (M), real split (M + 1)
  float2* tw_x = nullptr;
  float2* tw_y = nullptr;
  float2* tw_z = nullptr;
  float2* rt = nullptr;
  std::vector<std::pair<size_t, void*>> pool;  // retired member blocks kept for reuse
  int variant[4] = {0, 0, 0, 0};               // per tunable pass: 0 y, 1 x, 2 zu, 3 zp
  Profiler prof;
};

namespace {

int fill_twiddles(std::vector<float2>& out, int L) {
  for (int e = 0; e < L; ++e) {
    const double ang = 2.0 * M_PI * (double)e / (double)L;
    out.push_back(make_float2((float)std::cos(ang), (float)std::sin(ang)));
  }
  return L;
}

int check_slot(cst_session* s, int slot) {
  if (!s) return fail(CST_ERR_ARG, "null session");
  if (slot < 0 || slot >= s->max_members || !s->members[slot].used)
    return fail(CST_ERR_STATE, "member slot " + std::to_string(slot) + " is empty");
  return CST_OK;
}

// Drive amplitude float32(amplitude * env(n dt)) with env = (1 - cos(pi min(t / (ramp T),
// 1))) / 2 (sources.py:98-102), in float64, for the steps before the ramp completes; the
// device takes float32(amplitude) from then on (env = 1 exactly).
std::vector<float> ramp_table(double amplitude, double dt, double f0, double ramp_periods) {
  std::vector<float> out;
  const double period = 1.0 / f0;
  for (int64_t n = 0;; ++n) {
    const double x = std::min((double)n * dt / (ramp_periods * period), 1.0);
    if (x >= 1.0) break;
    out.push_back((float)(amplitude * (0.5 * (1.0 - std::cos(M_PI * x)))));
  }
  return out;
}

// (cos, sin)(2 pi k / spp) for k < spp in float64, rounded to float32: the drive phase
// 2 pi (n mod spp) / spp and every harmonic's record phase 2 pi (h n mod spp) / spp.
std::vector<float2> trig_table(int spp) {
  std::vector<float2> out(spp);
  for (int k = 0; k < spp; ++k) {
    const double th = 2.0 * M_PI * (double)k / (double)spp;
    out[k] = make_float2((float)std::cos(th), (float)std::sin(th));
  }
  return out;
}

int set_clock(cst_session* s, Member& m, int64_t n, int64_t record_from, int which) {
  hipLaunchKernelGGL(k_clock_set, dim3(1), dim3(64), 0, s->stream, m.dev.clock, (long long)n,
                     (long long)record_from, which);
  HIPCHK(hipGetLastError());
  return CST_OK;
}

bool batch_nonlinear(const cst_session* s, const StepArgs& a) {
  for (int i = 0; i < a.nmem; ++i)
    if (s->members[a.ids[i]].dev.beta2dt != nullptr) return true;
  return false;
}

bool batch_absorbing(const cst_session* s, const StepArgs& a) {
  return a.nmem > 0 && s->members[a.ids[0]].dev.stg.absorbing != 0;
}

// One step of the opt-in staggered scheme (the launch sequence of staggered_kernels.h).
// Timing kinds as the reference scheme's: 0 the gradient x pass, 1 the derivative y
// passes, 2 the velocity z pass, 3 the divergence and absorption x passes, 4 the plain y
// passes, 5 the density and pressure z passes.
int launch_step_staggered(cst_session* s, StepArgs& a) {
  const Shape& sh = s->s;
  const int M = sh.nz / 2;
  hipStream_t st = s->stream;
  const bool absorbing = batch_absorbing(s, a);
  bool ok = true;
  ok = ok && s->prof.run(0, st, [&] { return launch_sxgrad(sh.nx, a, s->tw_x, st); });
  ok = ok && s->prof.run(1, st, [&] { return launch_sygrad(sh.ny, a, s->tw_y, st); });
  ok = ok && s->prof.run(2, st, [&] { return launch_szu(M, a, s->tw_z, s->rt, st); });
  ok = ok && s->prof.run(1, st, [&] { return launch_sydiv(sh.ny, a, s->tw_y, st); });
  ok = ok && s->prof.run(3, st, [&] { return launch_sxdiv(sh.nx, a, s->tw_x, st); });
  if (sh.ny > 1)
    ok = ok && s->prof.run(4, st, [&] { return launch_y(sh.ny, +1, a, s->tw_y, 1, 0, st, 3); });
  ok = ok && s->prof.run(5, st, [&] { return launch_szrho(M, a, s->tw_z, s->rt, absorbing, st); });
  if (absorbing) {
    if (sh.ny > 1)
      ok = ok && s->prof.run(4, st, [&] { return launch_y(sh.ny, -1, a, s->tw_y, 1, 0, st, 2); });
    ok = ok && s->prof.run(3, st, [&] { return launch_sxabs(sh.nx, a, s->tw_x, st); });
    if (sh.ny > 1)
      ok = ok && s->prof.run(4, st, [&] { return launch_y(sh.ny, +1, a, s->tw_y, 1, 0, st, 2); });
    ok = ok && s->prof.run(5, st, [&] { return launch_szabs(M, a, s->tw_z, s->rt, st); });
  }
  if (sh.ny > 1)
    ok = ok && s->prof.run(4, st, [&] { return launch_y(sh.ny, -1, a, s->tw_y, 0, 0, st); });
  if (!ok) return fail(CST_ERR_UNSUPPORTED, "no kernel instantiated for this shape");
  HIPCHK(hipGetLastError());
  return CST_OK;
}

int launch_step(cst_session* s, StepArgs& a) {
  if (s->scheme == CST_SCHEME_STAGGERED) return launch_step_staggered(s, a);
  const Shape& sh = s->s;
  const int M = sh.nz / 2;
  hipStream_t st = s->stream;
  bool ok = true;
  ok = ok && s->prof.run(0, st, [&] { return launch_xgrad(sh.nx, s->variant[1], a, s->tw_x, st); });
  // the 3-spectrum y passes also apply the y/z derivative factors: they run for ny == 1 too
  ok = ok && s->prof.run(1, st, [&] { return launch_ygrad(sh.ny, s->variant[0], a, s->tw_y, st); });
  ok = ok && s->prof.run(2, st, [&] { return launch_zu(M, s->variant[2], a, s->tw_z, s->rt, st); });
  ok = ok && s->prof.run(1, st, [&] { return launch_ydiv(sh.ny, s->variant[0], a, s->tw_y, st); });
  ok = ok && s->prof.run(3, st, [&] { return launch_xdiv(sh.nx, s->variant[1], a, s->tw_x, st); });
  if (sh.ny > 1)
    ok = ok && s->prof.run(4, st, [&] { return launch_y(sh.ny, +1, a, s->tw_y, 0, s->variant[0], st); });
  const bool nonlinear = batch_nonlinear(s, a);
  ok = ok && s->prof.run(5, st, [&] {
    return launch_zp(M, s->variant[3], a, s->tw_z, s->rt, 1, nonlinear, st);
  });
  if (sh.ny > 1)
    ok = ok && s->prof.run(4, st, [&] { return launch_y(sh.ny, -1, a, s->tw_y, 0, s->variant[0], st); });
  if (!ok) return fail(CST_ERR_UNSUPPORTED, "no kernel instantiated for this shape");
  HIPCHK(hipGetLastError());
  return CST_OK;
}

// A new state restarts the record window: the sums and p_max are cleared (checkpoints
// are taken before the first record step, engine.py:344-371) and the clock set to n.
int restart_record(cst_session* s, Member& m, int64_t n) {
  HIPCHK(hipStreamSynchronize(s->io));
  for (int h = 0; h < m.nh; ++h) {
    HIPCHK(hipMemsetAsync(m.dev.acc[h], 0, m.rec_count * 8, s->stream));
    HIPCHK(hipMemsetAsync(m.dev.comp[h], 0, m.rec_count * 8, s->stream));
    m.finalized[h] = false;
  }
  HIPCHK(hipMemsetAsync(m.dev.pmax, 0, m.rec_count * 4, s->stream));
  return set_clock(s, m, n, 0, 1);
}

StepArgs single_member(const cst_session* s, int slot) {
  StepArgs a{};
  a.mem = s->d_table;
  a.s = s->s;
  a.nmem = 1;
  a.ids[0] = slot;
  return a;
}

void destroy_graphs(cst_session* s) {
  if (s->gcache.empty()) return;
  (void)hipStreamSynchronize(s->stream);
  for (auto& g : s->gcache) (void)hipGraphExecDestroy(g.exec);
  s->gcache.clear();
}

constexpr int kGraphSteps = 16;   // steps per replayed graph (a remainder replays 1-step graphs)
constexpr size_t kGraphCache = 48;

// The cached executable graph of `steps` steps for this member list and variant set,
// captured on first use; nullptr (and graphs off for the session) if capture fails.
hipGraphExec_t graph_for(cst_session* s, StepArgs& a, int steps) {
  const bool nl = batch_nonlinear(s, a);
  const bool ab = batch_absorbing(s, a);
  for (auto& g : s->gcache) {
    if (g.nmem != a.nmem || g.steps != steps || !std::equal(s->variant, s->variant + 4, g.v) ||
        g.nonlinear != nl || g.absorbing != ab)
      continue;
    if (!std::equal(a.ids, a.ids + a.nmem, g.ids)) continue;
    g.last_use = ++s->gclock;
    return g.exec;
  }
  if (s->gcache.size() >= kGraphCache) {
    auto it = std::min_element(s->gcache.begin(), s->gcache.end(),
                               [](const GraphEntry& x, const GraphEntry& y) { return x.last_use < y.last_use; });
    (void)hipStreamSynchronize(s->stream);  // the evicted graph may still be running
    (void)hipGraphExecDestroy(it->exec);
    s->gcache.erase(it);
  }
  hipGraph_t graph = nullptr;
  if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    s->graphs = false;
    return nullptr;
  }
  int r = CST_OK;
  for (int k = 0; k < steps && r == CST_OK; ++k) r = launch_step(s, a);
  const hipError_t e = hipStreamEndCapture(s->stream, &graph);
  hipGraphExec_t exec = nullptr;
  bool ok = r == CST_OK && e == hipSuccess && graph != nullptr &&
            hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
  if (graph) (void)hipGraphDestroy(graph);
  if (!ok) {
    (void)hipGetLastError();
    s->graphs = false;
    return nullptr;
  }
  GraphEntry g;
  g.nmem = a.nmem;
  std::copy(a.ids, a.ids + a.nmem, g.ids);
  std::copy(s->variant, s->variant + 4, g.v);
  g.nonlinear = nl;
  g.absorbing = ab;
  g.steps = steps;
  g.exec = exec;
  g.last_use = ++s->gclock;
  s->gcache.push_back(g);
  s->graph_captures++;
  return exec;
}

// R2C + forward y of the member's current p into sp (after set_state).
int prime(cst_session* s, int slot) {
  StepArgs a = single_member(s, slot);
  const int M = s->s.nz / 2;
  if (!launch_zp(M, s->variant[3], a, s->tw_z, s->rt, 0, false, s->stream))
    return fail(CST_ERR_UNSUPPORTED, "no kernel for nz");
  if (s->s.ny > 1 && !launch_y(s->s.ny, -1, a, s->tw_y, 0, s->variant[0], s->stream))
    return fail(CST_ERR_UNSUPPORTED, "no kernel for ny");
  HIPCHK(hipGetLastError());
  s->members[slot].primed = true;
  return CST_OK;
}

}  // namespace

namespace cst {
// plain float4 copy: the calibration kernel of cst_diag_copy
__global__ void __launch_bounds__(256) k_diag_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                                   size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
}  // namespace cst

extern "C" {

int cst_version(void) { return 103; }

int cst_last_error(char* buf, size_t n) { return copy_last_error(buf, n); }

int cst_device_count(int* n) {
  if (!n) return fail(CST_ERR_ARG, "null output");
  HIPCHK(hipGetDeviceCount(n));
  return CST_OK;
}

int cst_supported_length(int n, int axis) {
  if (n < 1) return 0;
  if (axis == 2) return (n % 2 == 0 && length_supported(n / 2)) ? 1 : 0;
  return length_supported(n) ? 1 : 0;
}

int cst_session_create(int nx, int ny, int nz, int max_members, int device, cst_session** out) {
  return cst_session_create_scheme(nx, ny, nz, max_members, device, CST_SCHEME_REFERENCE, out);
}

int cst_session_scheme(cst_session* s, int* scheme) {
  if (!s || !scheme) return fail(CST_ERR_ARG, "null argument");
  *scheme = s->scheme;
  return CST_OK;
}

int cst_session_create_scheme(int nx, int ny, int nz, int max_members, int device, int scheme,
                              cst_session** out) {
  if (!out) return fail(CST_ERR_ARG, "null output");
  *out = nullptr;
  if (scheme != CST_SCHEME_REFERENCE && scheme != CST_SCHEME_STAGGERED)
    return fail(CST_ERR_ARG, "scheme must be CST_SCHEME_REFERENCE (0) or CST_SCHEME_STAGGERED (1)");
  if (nx < 1 || ny < 1 || nz < 2 || max_members < 1 || max_members > 4096)
    return fail(CST_ERR_ARG, "invalid shape or member count");
  if (!cst_supported_length(nx, 0) || !cst_supported_length(ny, 1) || !cst_supported_length(nz, 2))
    return fail(CST_ERR_UNSUPPORTED, "padded shape (" + std::to_string(nx) + ", " + std::to_string(ny) +
                                         ", " + std::to_string(nz) +
                                         ") has an axis without an instantiated kernel (z must be even)");
  HIPCHK(hipSetDevice(device));
  auto* s = new cst_session();
  s->device = device;
  s->scheme = scheme;
  {
    const int ms = ((nz / 2 + 15) / 16) * 16;
    const long ps = (long)ny * ms + 16, so = (long)nx * ps;
    s->s = Shape{nx, ny, nz, nz / 2 + 1, ms, ps, so, so + (((long)nx * ny + 15) / 16) * 16};
  }
  s->max_members = max_members;
  s->members.resize(max_members);
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->io, hipStreamNonBlocking);
  if (e != hipSuccess) {
    cst_session_destroy(s);
    return fail(CST_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  std::vector<float2> tab;
  fill_twiddles(tab, nx);
  fill_twiddles(tab, ny);
  fill_twiddles(tab, nz / 2);
  for (int k = 0; k <= nz / 2; ++k) {
    const double ang = 2.0 * M_PI * (double)k / (double)nz;
    tab.push_back(make_float2((float)std::cos(ang), (float)std::sin(ang)));
  }
  e = hipMalloc(&s->d_tables, tab.size() * sizeof(float2));
  if (e == hipSuccess) e = hipMalloc(&s->d_table, sizeof(MemberDev) * max_members);
  if (e == hipSuccess)
    e = hipMemcpy(s->d_tables, tab.data(), tab.size() * sizeof(float2), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(s->d_table, 0, sizeof(MemberDev) * max_members);
  if (e != hipSuccess) {
    cst_session_destroy(s);
    return fail(CST_ERR_HIP, std::string("session tables: ") + hipGetErrorString(e));
  }
  s->tw_x = s->d_tables;
  s->tw_y = s->tw_x + nx;
  s->tw_z = s->tw_y + ny;
  s->rt = s->tw_z + nz / 2;
  *out = s;
  return CST_OK;
}

int cst_session_destroy(cst_session* s) {
  if (!s) return CST_OK;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->io) (void)hipStreamSynchronize(s->io);
  if (s->up) (void)hipStreamSynchronize(s->up);
  destroy_graphs(s);
  for (auto& m : s->members) {
    if (m.block) (void)hipFree(m.block);
    if (m.done) (void)hipEventDestroy(m.done);
  }
  for (auto& b : s->pool) (void)hipFree(b.second);
  if (s->d_table) (void)hipFree(s->d_table);
  if (s->d_tables) (void)hipFree(s->d_tables);
  s->prof.release();
  if (s->stream) (void)hipStreamDestroy(s->stream);
  if (s->up) (void)hipStreamDestroy(s->up);
  if (s->io) (void)hipStreamDestroy(s->io);
  delete s;
  return CST_OK;
}

int cst_session_stream(cst_session* s, void** stream) {
  if (!s || !stream) return fail(CST_ERR_ARG, "null argument");
  *stream = (void*)s->stream;
  return CST_OK;
}

}  // extern "C"

namespace {

// copy kind for a checked tensor as source or destination of a device-side buffer
hipMemcpyKind to_device(const dl::View& v) {
  return v.where == dl::Where::Device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
}
hipMemcpyKind from_device(const dl::View& v) {
  return v.where == dl::Where::Device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
}

int check_dl(const DLTensor* t, const char* what, DLDataType want, int64_t count, int device,
             dl::View* out) {
  std::string err;
  if (!dl::check(t, what, want, count, nullptr, 0, device, out, &err)) return fail(CST_ERR_ARG, err);
  return CST_OK;
}

// Staggered-scheme inputs of a member (cst_staggered_desc and its nine maps).
struct StgIn {
  const cst_staggered_desc* sd;
  dl::View maps[9];
};
constexpr int kStgMaps = 9;

int pml_stride(const Shape& sh) { return ((sh.nz + sh.nx + sh.ny + 63) / 64) * 64; }

// maps[0..3]: dt/rho, rho c^2 dt, absorb, 2 beta dt (data nullptr: linear); stg: the
// staggered scheme's inputs (nullptr for a reference-scheme member, whose maps they replace)
int member_add(cst_session* s, const cst_member_desc* d, const dl::View maps[4], const StgIn* stg,
               int* slot_out) {
  const Shape& sh = s->s;
  if ((stg != nullptr) != (s->scheme == CST_SCHEME_STAGGERED))
    return fail(CST_ERR_ARG, s->scheme == CST_SCHEME_STAGGERED
                                 ? "a staggered session takes members through cst_member_add_staggered*"
                                 : "cst_member_add_staggered* needs a session of scheme CST_SCHEME_STAGGERED");
  if (d->n_harmonics < 1 || d->n_harmonics > kMaxHarmonics)
    return fail(CST_ERR_ARG, "n_harmonics must be 1..3");
  if (d->spp < 2 || !(d->dt > 0) || !(d->f0 > 0) || !(d->amplitude > 0) || !(d->ramp_periods > 0))
    return fail(CST_ERR_ARG, "invalid timing (spp, dt, f0, amplitude, ramp)");
  if (!stg && (!maps[0].data || !maps[1].data || !maps[2].data || !d->ktab || !d->sponge))
    return fail(CST_ERR_ARG, "missing coefficient maps, k tables or sponge profiles");
  bool absorbing = false, stg_nl = false;
  if (stg) {
    const cst_staggered_desc* sd = stg->sd;
    absorbing = sd->absorbing != 0;
    stg_nl = stg->maps[7].data != nullptr;
    for (int i = 0; i < 5; ++i)
      if (!stg->maps[i].data) return fail(CST_ERR_ARG, "staggered member: maps 0..4 are required");
    if (absorbing && (!stg->maps[5].data || !stg->maps[6].data))
      return fail(CST_ERR_ARG, "absorbing staggered member: maps 5 (c^2 tau / dt) and 6 (c^2 eta) are required");
    if (stg_nl != (stg->maps[8].data != nullptr))
      return fail(CST_ERR_ARG, "nonlinear staggered member: maps 7 (2 m dt) and 8 (c^2 m (beta - 1) / rho0) go together");
    if (!sd->kfac || !sd->pml || !d->ktab) return fail(CST_ERR_ARG, "missing k tables or PML tables");
    if (absorbing && !(sd->alpha_power > 0.0 && sd->alpha_power < 3.0 && std::fabs(sd->alpha_power - 1.0) > 1e-6))
      return fail(CST_ERR_ARG, "alpha_power must lie in (0, 3) and differ from 1");
    if (sd->src_axes < 0 || sd->src_axes > 7) return fail(CST_ERR_ARG, "src_axes is a 3-bit axis mask");
  }
  if (d->n_src < 0 || (d->n_src > 0 && (!d->src_idx || !d->src_phase || !d->src_scale)))
    return fail(CST_ERR_ARG, "invalid source arrays");
  const int dims[3] = {sh.nx, sh.ny, sh.nz};
  for (int ax = 0; ax < 3; ++ax) {
    const int c0 = d->conv[2 * ax], c1 = d->conv[2 * ax + 1];
    const int r0 = d->rec[2 * ax], r1 = d->rec[2 * ax + 1];
    if (c0 < 0 || c1 > dims[ax] || c0 >= c1 || r0 < 0 || r1 > dims[ax] || r0 >= r1)
      return fail(CST_ERR_ARG, "convergence or record box outside the padded grid");
  }
  int slot = -1;
  for (int i = 0; i < s->max_members; ++i)
    if (!s->members[i].used) {
      slot = i;
      break;
    }
  if (slot < 0) return fail(CST_ERR_STATE, "session is full");

  // sources: CSR over (x, y) lines, duplicates refused
  const int nl = sh.nx * sh.ny;
  std::vector<int> ptr(nl + 1, 0), order(d->n_src);
  for (int j = 0; j < d->n_src; ++j) {
    const int32_t* v = d->src_idx + 3 * j;
    if (v[0] < 0 || v[0] >= sh.nx || v[1] < 0 || v[1] >= sh.ny || v[2] < 0 || v[2] >= sh.nz)
      return fail(CST_ERR_ARG, "source voxel outside the padded grid");
    order[j] = j;
  }
  auto key = [&](int j) {
    const int32_t* v = d->src_idx + 3 * j;
    return ((int64_t)v[0] * sh.ny + v[1]) * sh.nz + v[2];
  };
  std::sort(order.begin(), order.end(), [&](int a, int b) { return key(a) < key(b); });
  for (int j = 1; j < d->n_src; ++j)
    if (key(order[j]) == key(order[j - 1])) return fail(CST_ERR_ARG, "duplicate source voxel");
  std::vector<int> src_z(std::max(d->n_src, 1));
  std::vector<float2> src_ab(std::max(d->n_src, 1));
  for (int j = 0; j < d->n_src; ++j) {
    const int o = order[j];
    const int32_t* v = d->src_idx + 3 * o;
    ptr[v[0] * sh.ny + v[1] + 1]++;
    src_z[j] = v[2];
    const double sc = d->src_scale[o], ph = d->src_phase[o];
    src_ab[j] = make_float2((float)(sc * std::cos(ph)), (float)(sc * std::sin(ph)));
  }
  for (int i = 0; i < nl; ++i) ptr[i + 1] += ptr[i];

  // one device allocation, carved in 256-byte aligned pieces
  const int64_t P = (int64_t)sh.nx * sh.ny * sh.nz;
  const int64_t H = sh.hs;  // bins per spectrum array (rows + Nyquist side plane)
  const int64_t R = (int64_t)(d->rec[1] - d->rec[0]) * (d->rec[3] - d->rec[2]) * (d->rec[5] - d->rec[4]);
  const bool nonlinear = stg ? false : maps[3].data != nullptr;
  const int nmaps = stg ? 5 + (absorbing ? 2 : 0) + (stg_nl ? 2 : 0) : (nonlinear ? 4 : 3);
  const int64_t nk = 2LL * (sh.nx + sh.ny + sh.nzh);
  const int64_t nkf = (int64_t)sh.nx + sh.ny + sh.nzh;  // staggered shift factors (float2)
  const int pst = pml_stride(sh);
  const int64_t nsp = sh.nz + sh.nx + sh.ny;
  const int maxp = std::max(d->max_periods, 1);
  const std::vector<float> amp_tab = ramp_table(d->amplitude, d->dt, d->f0, d->ramp_periods);
  const std::vector<float2> trig = trig_table(d->spp);
  const long long clock0[2] = {0, INT64_MAX};
  Carver c;
  const size_t o_state = c.take(P * 4);      // p
  const size_t o_spec = c.take(7 * H * 8);   // sp, sg0..sg2, su0..su2 (velocity state)
  const size_t o_rho = stg ? c.take(3 * P * 4) : 0;  // split density (staggered)
  const size_t o_maps = c.take(nmaps * P * 4);
  const size_t o_k = c.take(nk * 4);
  // reference: sponge profiles; staggered: 4 PML tables, then the shift factors
  const size_t o_sp = stg ? c.take(4LL * pst * 4) : c.take(nsp * 4);
  const size_t o_kf = stg ? c.take(nkf * 8) : 0;
  const size_t o_ptr = c.take((nl + 1) * 4);
  // at least one entry each: the z pass loads index 0 unconditionally
  const size_t o_sz = c.take((int64_t)std::max<size_t>(src_z.size(), 1) * 4);
  const size_t o_sab = c.take((int64_t)std::max<size_t>(src_ab.size(), 1) * 8);
  const size_t o_rec = c.take(2 * d->n_harmonics * R * 8);
  const size_t o_pmax = c.take(R * 4);
  const size_t o_peaks = c.take((int64_t)maxp * kPeakWords * 4);
  // the clock (16 B) and, 64 B on, the step scalars k_xgrad stores: one 256-B slot
  constexpr size_t kStepvAt = 64;
  static_assert(kStepvAt >= sizeof(clock0) && kStepvAt + sizeof(StepMember) <= 256, "one 256-B slot");
  const size_t o_clock = c.take(kStepvAt + sizeof(StepMember));
  const size_t o_amp = c.take((int64_t)std::max<size_t>(amp_tab.size(), 1) * 4);
  const size_t o_trig = c.take((int64_t)trig.size() * 8);
  HIPCHK(hipSetDevice(s->device));
  Member& m = s->members[slot];
  char* base = nullptr;
  for (size_t i = 0; i < s->pool.size(); ++i)
    if (s->pool[i].first == c.total) {
      base = (char*)s->pool[i].second;
      s->pool.erase(s->pool.begin() + i);
      break;
    }
  if (!base) HIPCHK(hipMalloc((void**)&base, c.total));
  hipEvent_t done = m.done;  // the slot's event outlives its members
  if (!done) {
    const hipError_t ee = hipEventCreateWithFlags(&done, hipEventDisableTiming);
    if (ee != hipSuccess) {
      s->pool.emplace_back(c.total, base);
      return fail(CST_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(ee));
    }
  }
  m = Member{};
  m.done = done;
  m.block = base;
  m.block_bytes = c.total;
  m.used = true;
  MemberDev& md = m.dev;
  md.p = (float*)(base + o_state);
  float2* spc = (float2*)(base + o_spec);
  md.sp = spc;
  for (int a = 0; a < 3; ++a) md.sg[a] = spc + (a + 1) * H;
  for (int a = 0; a < 3; ++a) md.su[a] = spc + (a + 4) * H;
  float* maps_d = (float*)(base + o_maps);
  if (stg) {
    // the reference-scheme maps are never read on a staggered session
    md.dto_rho = md.rhoc2dt = md.absorb = md.beta2dt = nullptr;
    StgDev& g = md.stg;
    for (int ax = 0; ax < 3; ++ax) {
      g.rho[ax] = (float*)(base + o_rho) + ax * P;
      g.dtrho[ax] = maps_d + ax * P;
    }
    g.dtrho0 = maps_d + 3 * P;
    g.c2 = maps_d + 4 * P;
    int k = 5;
    g.c2tau = absorbing ? maps_d + (k++) * P : nullptr;
    g.c2eta = absorbing ? maps_d + (k++) * P : nullptr;
    g.nl2dt = stg_nl ? maps_d + (k++) * P : nullptr;
    g.c2bh = stg_nl ? maps_d + (k++) * P : nullptr;
    g.kfac = (const float2*)(base + o_kf);
    g.pml = (const float*)(base + o_sp);
    g.pml_stride = pst;
    g.e1 = (float)((stg->sd->alpha_power - 2.0) / 2.0);
    g.e2 = (float)((stg->sd->alpha_power - 1.0) / 2.0);
    g.absorbing = absorbing ? 1 : 0;
    g.src_axes = stg->sd->src_axes;
  } else {
    md.dto_rho = maps_d;
    md.rhoc2dt = maps_d + P;
    md.absorb = maps_d + 2 * P;
    md.beta2dt = nonlinear ? maps_d + 3 * P : nullptr;
  }
  md.ktab = (float*)(base + o_k);
  // staggered: the grid-point PML table (read by the plain R2C pass of prime())
  md.sponge = (float*)(base + o_sp);
  md.src_ptr = (int*)(base + o_ptr);
  md.src_z = (int*)(base + o_sz);
  md.src_ab = (float2*)(base + o_sab);
  float2* rec = (float2*)(base + o_rec);
  for (int h = 0; h < kMaxHarmonics; ++h) {
    md.acc[h] = h < d->n_harmonics ? rec + (2 * h) * R : nullptr;
    md.comp[h] = h < d->n_harmonics ? rec + (2 * h + 1) * R : nullptr;
  }
  md.pmax = (float*)(base + o_pmax);
  md.peaks = (unsigned*)(base + o_peaks);
  for (int i = 0; i < 6; ++i) {
    md.conv[i] = d->conv[i];
    md.rec[i] = d->rec[i];
  }
  md.nh = d->n_harmonics;
  md.clock = (long long*)(base + o_clock);
  md.stepv = (StepMember*)(base + o_clock + kStepvAt);
  md.amp_tab = (const float*)(base + o_amp);
  md.trig = (const float2*)(base + o_trig);
  md.amp_full = (float)(d->amplitude * (0.5 * (1.0 - std::cos(M_PI * 1.0))));
  md.ramp_n = (int)amp_tab.size();
  md.spp = d->spp;
  md.max_periods = maxp;
  for (int h = 0; h < kMaxHarmonics; ++h) md.harm[h] = h < d->n_harmonics ? d->harmonics[h] : 0;
  md.half_cdt = (float)(0.5 * d->c_ref * d->dt);
  md.inv_n = (float)(1.0 / (double)P);
  md.inv_nz = (float)(1.0 / (double)sh.nz);

  m.dt = d->dt;
  m.f0 = d->f0;
  m.amplitude = d->amplitude;
  m.ramp_periods = d->ramp_periods;
  m.spp = d->spp;
  m.nh = d->n_harmonics;
  for (int h = 0; h < kMaxHarmonics; ++h) m.harmonics[h] = h < d->n_harmonics ? d->harmonics[h] : 0;
  m.max_periods = maxp;

  // uploads on the upload stream: members of the session may be running on the
  // compute stream meanwhile (the block and the table entry are this member's alone)
  hipStream_t stm = s->up;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stm);
  };
  hipError_t e = hipMemsetAsync(base + o_state, 0, P * 4, stm);
  if (e == hipSuccess) e = hipMemsetAsync(base + o_spec, 0, 7 * H * 8, stm);
  if (stg) {
    if (e == hipSuccess) e = hipMemsetAsync(base + o_rho, 0, 3 * P * 4, stm);
    int k = 0;
    for (int i = 0; i < kStgMaps && e == hipSuccess; ++i) {
      if (!stg->maps[i].data) continue;
      e = hipMemcpyAsync(maps_d + (k++) * P, stg->maps[i].data, P * 4, to_device(stg->maps[i]), stm);
    }
    for (int t = 0; t < 4 && e == hipSuccess; ++t)
      e = up((char*)(base + o_sp) + (size_t)t * pst * 4, stg->sd->pml + (size_t)t * nsp, nsp * 4);
    if (e == hipSuccess) e = up(base + o_kf, stg->sd->kfac, nkf * 8);
  } else {
    // the maps come from host or device memory (DLPack views checked by the caller)
    float* const dst_map[4] = {(float*)md.dto_rho, (float*)md.rhoc2dt, (float*)md.absorb, (float*)md.beta2dt};
    for (int i = 0; i < nmaps && e == hipSuccess; ++i)
      e = hipMemcpyAsync(dst_map[i], maps[i].data, P * 4, to_device(maps[i]), stm);
    if (e == hipSuccess) e = up((void*)md.sponge, d->sponge, nsp * 4);
  }
  if (e == hipSuccess) e = up((void*)md.ktab, d->ktab, nk * 4);
  if (e == hipSuccess) e = up((void*)md.src_ptr, ptr.data(), (nl + 1) * 4);
  if (e == hipSuccess) e = up((void*)md.src_z, src_z.data(), src_z.size() * 4);
  if (e == hipSuccess) e = up((void*)md.src_ab, src_ab.data(), src_ab.size() * 8);
  if (e == hipSuccess) e = hipMemsetAsync(base + o_rec, 0, 2 * d->n_harmonics * R * 8, stm);
  if (e == hipSuccess) e = hipMemsetAsync(base + o_pmax, 0, R * 4, stm);
  if (e == hipSuccess) e = hipMemsetAsync(md.peaks, 0, (size_t)maxp * kPeakWords * 4, stm);
  if (e == hipSuccess) e = up(md.clock, clock0, sizeof(clock0));
  if (e == hipSuccess && !amp_tab.empty()) e = up((void*)md.amp_tab, amp_tab.data(), amp_tab.size() * 4);
  if (e == hipSuccess) e = up((void*)md.trig, trig.data(), trig.size() * 8);
  if (e == hipSuccess) e = up(s->d_table + slot, &md, sizeof(MemberDev));
  if (e == hipSuccess) e = hipStreamSynchronize(stm);
  if (e != hipSuccess) {
    s->pool.emplace_back(c.total, base);
    m = Member{};
    m.done = done;
    return fail(CST_ERR_HIP, std::string("member upload: ") + hipGetErrorString(e));
  }
  m.rec_count = R;
  *slot_out = slot;
  return CST_OK;
}

}  // namespace

extern "C" {

int cst_member_add_dl(cst_session* s, const cst_member_desc* d, const DLTensor* dt_over_rho,
                      const DLTensor* rhoc2_dt, const DLTensor* absorb, const DLTensor* beta2_dt,
                      int* slot) {
  if (!s || !d || !slot) return fail(CST_ERR_ARG, "null argument");
  const int64_t P = (int64_t)s->s.nx * s->s.ny * s->s.nz;
  dl::View v[4];
  const DLTensor* t[4] = {dt_over_rho, rhoc2_dt, absorb, beta2_dt};
  const char* names[4] = {"dt_over_rho", "rhoc2_dt", "absorb", "beta2_dt"};
  for (int i = 0; i < 4; ++i) {
    if (i == 3 && !beta2_dt) break;
    if (int r = check_dl(t[i], names[i], dl::kF32, P, s->device, &v[i])) return r;
  }
  return member_add(s, d, v, nullptr, slot);
}

int cst_member_add_staggered_dl(cst_session* s, const cst_member_desc* d, const cst_staggered_desc* sd,
                                const DLTensor* const* maps, int* slot) {
  if (!s || !d || !sd || !maps || !slot) return fail(CST_ERR_ARG, "null argument");
  const int64_t P = (int64_t)s->s.nx * s->s.ny * s->s.nz;
  StgIn in{};
  in.sd = sd;
  static const char* names[kStgMaps] = {"dtrho_x", "dtrho_y", "dtrho_z", "dtrho0", "c2",
                                        "c2tau", "c2eta", "nl2dt", "c2bh"};
  for (int i = 0; i < kStgMaps; ++i) {
    if (!maps[i]) continue;
    if (int r = check_dl(maps[i], names[i], dl::kF32, P, s->device, &in.maps[i])) return r;
  }
  dl::View none[4];
  return member_add(s, d, none, &in, slot);
}

int cst_member_add_staggered(cst_session* s, const cst_member_desc* d, const cst_staggered_desc* sd,
                             const float* const* maps, int* slot) {
  if (!s || !d || !sd || !maps || !slot) return fail(CST_ERR_ARG, "null argument");
  int64_t shp[3] = {s->s.nx, s->s.ny, s->s.nz};
  DLTensor t[kStgMaps];
  const DLTensor* tp[kStgMaps];
  for (int i = 0; i < kStgMaps; ++i) {
    t[i] = dl::host_tensor(const_cast<float*>(maps[i]), dl::kF32, shp, 3);
    tp[i] = maps[i] ? &t[i] : nullptr;
  }
  return cst_member_add_staggered_dl(s, d, sd, tp, slot);
}

int cst_member_add(cst_session* s, const cst_member_desc* d, int* slot_out) {
  if (!s || !d || !slot_out) return fail(CST_ERR_ARG, "null argument");
  int64_t shp[3] = {s->s.nx, s->s.ny, s->s.nz};
  const float* src[4] = {d->dt_over_rho, d->rhoc2_dt, d->absorb, d->beta2_dt};
  DLTensor t[4];
  for (int i = 0; i < 4; ++i) t[i] = dl::host_tensor(const_cast<float*>(src[i]), dl::kF32, shp, 3);
  if (!d->dt_over_rho || !d->rhoc2_dt || !d->absorb)
    return fail(CST_ERR_ARG, "missing coefficient maps, k tables or sponge profiles");
  return cst_member_add_dl(s, d, &t[0], &t[1], &t[2], d->beta2_dt ? &t[3] : nullptr, slot_out);
}

int cst_member_remove(cst_session* s, int slot) {
  if (int r = check_slot(s, slot)) return r;
  HIPCHK(hipSetDevice(s->device));
  // the member's own last step, not the whole stream: other members' steps queued after
  // it (a chunk launched before this retirement) need not finish first
  HIPCHK(hipEventSynchronize(s->members[slot].done));
  HIPCHK(hipStreamSynchronize(s->io));
  // keep the block for the next member of this shape (hipFree would stall every stream)
  s->pool.emplace_back(s->members[slot].block_bytes, s->members[slot].block);
  hipEvent_t done = s->members[slot].done;
  s->members[slot] = Member{};
  s->members[slot].done = done;
  return CST_OK;
}

int cst_session_trim(cst_session* s, size_t* released) {
  if (!s) return fail(CST_ERR_ARG, "null session");
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipStreamSynchronize(s->stream));
  size_t n = 0;
  for (auto& b : s->pool) {
    n += b.first;
    HIPCHK(hipFree(b.second));
  }
  s->pool.clear();
  if (released) *released = n;
  return CST_OK;
}

int cst_session_pooled_bytes(cst_session* s, size_t* bytes) {
  if (!s || !bytes) return fail(CST_ERR_ARG, "null argument");
  size_t n = 0;
  for (auto& b : s->pool) n += b.first;
  *bytes = n;
  return CST_OK;
}

int cst_diag_copy(int device, size_t bytes) {
  if (bytes == 0 || bytes % 16 != 0 || bytes > ((size_t)1 << 34))
    return fail(CST_ERR_ARG, "bytes must be a positive multiple of 16, at most 16 GiB");
  HIPCHK(hipSetDevice(device));
  void* a = nullptr;
  void* b = nullptr;
  HIPCHK(hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return fail(CST_ERR_HIP, "hipMalloc failed for the copy target");
  }
  const size_t n = bytes / 16;
  hipLaunchKernelGGL(cst::k_diag_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr,
                     static_cast<const float4*>(a), static_cast<float4*>(b), n);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return fail(CST_ERR_HIP, hipGetErrorString(e));
  return CST_OK;
}

int cst_device_memory(int device, size_t* free_bytes, size_t* total_bytes) {
  if (!free_bytes || !total_bytes) return fail(CST_ERR_ARG, "null output");
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipMemGetInfo(free_bytes, total_bytes));
  return CST_OK;
}

int cst_member_reset(cst_session* s, int slot) {
  if (int r = check_slot(s, slot)) return r;
  Member& m = s->members[slot];
  const Shape& sh = s->s;
  const int64_t P = (int64_t)sh.nx * sh.ny * sh.nz;
  const int64_t H = sh.hs;
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipMemsetAsync(m.dev.p, 0, P * 4, s->stream));
  HIPCHK(hipMemsetAsync(m.dev.sp, 0, 7 * H * 8, s->stream));
  if (s->scheme == CST_SCHEME_STAGGERED)
    for (int ax = 0; ax < 3; ++ax) HIPCHK(hipMemsetAsync(m.dev.stg.rho[ax], 0, P * 4, s->stream));
  for (int h = 0; h < m.nh; ++h) {
    HIPCHK(hipMemsetAsync(m.dev.acc[h], 0, m.rec_count * 8, s->stream));
    HIPCHK(hipMemsetAsync(m.dev.comp[h], 0, m.rec_count * 8, s->stream));
    m.finalized[h] = false;
  }
  HIPCHK(hipMemsetAsync(m.dev.pmax, 0, m.rec_count * 4, s->stream));
  HIPCHK(hipMemsetAsync(m.dev.peaks, 0, (size_t)m.max_periods * kPeakWords * 4, s->stream));
  if (int r = set_clock(s, m, 0, INT64_MAX, 3)) return r;
  HIPCHK(hipEventRecord(m.done, s->stream));  // cst_member_remove waits for this work too
  m.n = 0;
  m.record_from = INT64_MAX;
  m.primed = true;
  return CST_OK;
}

int cst_member_schedule(cst_session* s, int slot, int64_t record_from) {
  if (int r = check_slot(s, slot)) return r;
  HIPCHK(hipSetDevice(s->device));
  if (int r = set_clock(s, s->members[slot], 0, record_from, 2)) return r;
  HIPCHK(hipEventRecord(s->members[slot].done, s->stream));
  s->members[slot].record_from = record_from;
  return CST_OK;
}

int cst_member_step(cst_session* s, int slot, int64_t* n) {
  if (int r = check_slot(s, slot)) return r;
  if (!n) return fail(CST_ERR_ARG, "null output");
  *n = s->members[slot].n;
  return CST_OK;
}

int cst_advance(cst_session* s, int n_slots, const int* slots, int n_steps) {
  if (!s || (n_slots > 0 && !slots)) return fail(CST_ERR_ARG, "null argument");
  if (n_slots < 1 || n_slots > kMaxBatch) return fail(CST_ERR_ARG, "1..16 members per advance");
  if (n_steps < 0) return fail(CST_ERR_ARG, "negative step count");
  for (int i = 0; i < n_slots; ++i) {
    if (int r = check_slot(s, slots[i])) return r;
    for (int j = 0; j < i; ++j)
      if (slots[j] == slots[i]) return fail(CST_ERR_ARG, "member listed twice");
  }
  for (int i = 0; i < n_slots; ++i) {
    const Member& m = s->members[slots[i]];
    for (int h = 0; h < m.nh; ++h)
      if (m.finalized[h])
        return fail(CST_ERR_STATE, "member " + std::to_string(slots[i]) +
                                       ": its phasors were finalized by cst_member_result; reset it "
                                       "(cst_member_reset) or give it a new state before advancing it");
  }
  for (int i = 1; i < n_slots; ++i)
    if (s->members[slots[i]].dev.stg.absorbing != s->members[slots[0]].dev.stg.absorbing)
      return fail(CST_ERR_ARG, "staggered members advanced together must all be absorbing or all not");
  HIPCHK(hipSetDevice(s->device));
  for (int i = 0; i < n_slots; ++i)
    if (!s->members[slots[i]].primed)
      if (int r = prime(s, slots[i])) return r;
  StepArgs a{};
  a.mem = s->d_table;
  a.s = s->s;
  a.nmem = n_slots;
  for (int i = 0; i < n_slots; ++i) a.ids[i] = slots[i];
  // every step of the sequence takes the same arguments (the step scalars come from the
  // members' device clocks): replay captured graphs unless the launches are being timed
  int k = 0;
  if (s->graphs && !s->prof.on && n_steps >= 2) {
    if (n_steps >= kGraphSteps) {
      if (hipGraphExec_t g = graph_for(s, a, kGraphSteps)) {
        for (; k + kGraphSteps <= n_steps; k += kGraphSteps) {
          HIPCHK(hipGraphLaunch(g, s->stream));
          s->graph_launches++;
        }
      }
    }
    if (s->graphs && k < n_steps) {
      if (hipGraphExec_t g = graph_for(s, a, 1)) {
        for (; k < n_steps; ++k) {
          HIPCHK(hipGraphLaunch(g, s->stream));
          s->graph_launches++;
        }
      }
    }
  }
  for (; k < n_steps; ++k)
    if (int r = launch_step(s, a)) return r;
  for (int i = 0; i < n_slots; ++i) {
    Member& m = s->members[slots[i]];
    m.n += n_steps;
    HIPCHK(hipEventRecord(m.done, s->stream));
  }
  return CST_OK;
}

int cst_sync(cst_session* s) {
  if (!s) return fail(CST_ERR_ARG, "null session");
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipStreamSynchronize(s->stream));
  return CST_OK;
}

int cst_session_query(cst_session* s, int* idle) {
  if (!s || !idle) return fail(CST_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(s->device));
  const hipError_t e = hipStreamQuery(s->stream);
  if (e == hipSuccess) {
    *idle = 1;
  } else if (e == hipErrorNotReady) {
    *idle = 0;
  } else {
    return fail(CST_ERR_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
  }
  return CST_OK;
}

int cst_member_peaks(cst_session* s, int slot, int first, int count, float* out) {
  if (int r = check_slot(s, slot)) return r;
  Member& m = s->members[slot];
  if (!out || first < 0 || count < 0 || first + count > m.max_periods)
    return fail(CST_ERR_ARG, "peak range outside the member's period capacity");
  if (count == 0) return CST_OK;
  HIPCHK(hipSetDevice(s->device));
  std::vector<unsigned> raw((size_t)count * kPeakWords);
  // a host wait for the member's last queued step, not a stream dependency: another thread
  // may be capturing the compute stream (the next steps of the other members) meanwhile,
  // and a dependency of the io stream on an event of a capturing stream is refused
  HIPCHK(hipEventSynchronize(m.done));
  HIPCHK(hipMemcpyAsync(raw.data(), m.dev.peaks + (size_t)first * kPeakWords,
                        raw.size() * 4, hipMemcpyDeviceToHost, s->io));
  HIPCHK(hipStreamSynchronize(s->io));
  for (int i = 0; i < count; ++i) {
    unsigned b = 0u;
    for (int k = 0; k < kPeakSlots; ++k) b = std::max(b, raw[(size_t)i * kPeakWords + k * kPeakStride]);
    std::memcpy(out + i, &b, 4);
  }
  return CST_OK;
}

int cst_member_result_into(cst_session* s, int slot, int hi, double scale, DLTensor* phasor,
                           DLTensor* pmax) {
  if (int r = check_slot(s, slot)) return r;
  Member& m = s->members[slot];
  if (hi < 0 || hi >= m.nh) return fail(CST_ERR_ARG, "harmonic index out of range");
  if (m.finalized[hi] && m.final_scale[hi] != scale)
    return fail(CST_ERR_ARG, "harmonic " + std::to_string(hi) +
                                 " was already finalized with another scale (the reduction runs once)");
  const int64_t R = m.rec_count;
  dl::View vp, vm;
  if (phasor)
    if (int r = check_dl(phasor, "phasor", dl::kC64, R, s->device, &vp)) return r;
  if (pmax)
    if (int r = check_dl(pmax, "pmax", dl::kF32, R, s->device, &vm)) return r;
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipEventSynchronize(m.done));  // a host wait, as in cst_member_peaks
  if (!m.finalized[hi]) {
    const int threads = 256;
    const int64_t blocks = (R + threads - 1) / threads;
    s->prof.run(6, s->io, [&] {
      hipLaunchKernelGGL(k_finalize, dim3((unsigned)blocks), dim3(threads), 0, s->io, m.dev.acc[hi],
                         m.dev.comp[hi], R, scale);
      return true;
    });
    HIPCHK(hipGetLastError());
    m.finalized[hi] = true;
    m.final_scale[hi] = scale;
  }
  if (phasor) HIPCHK(hipMemcpyAsync(vp.data, m.dev.comp[hi], R * 8, from_device(vp), s->io));
  if (pmax) HIPCHK(hipMemcpyAsync(vm.data, m.dev.pmax, R * 4, from_device(vm), s->io));
  HIPCHK(hipStreamSynchronize(s->io));
  return CST_OK;
}

int cst_member_result(cst_session* s, int slot, int hi, double scale, float* phasor, float* pmax) {
  if (int r = check_slot(s, slot)) return r;
  int64_t R = s->members[slot].rec_count;
  DLTensor tp = dl::host_tensor(phasor, dl::kC64, &R, 1);
  DLTensor tm = dl::host_tensor(pmax, dl::kF32, &R, 1);
  return cst_member_result_into(s, slot, hi, scale, phasor ? &tp : nullptr, pmax ? &tm : nullptr);
}

}  // extern "C"

namespace {

void release_export(void* data, dl::Where where, int device) {
  if (where == dl::Where::Device) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    (void)hipFree(data);
    (void)hipSetDevice(cur);
  } else {
    std::free(data);
  }
}

// a new managed tensor of `bytes` on the host (kDLCPU) or the session's GPU (kDLROCM)
int alloc_export(int device_type, int device, DLDataType dt, const int64_t* shape, int ndim,
                 DLManagedTensor** out) {
  int64_t n = 1;
  for (int i = 0; i < ndim; ++i) n *= shape[i];
  const size_t bytes = (size_t)std::max<int64_t>(n, 1) * (dt.bits / 8) * dt.lanes;
  void* data = nullptr;
  dl::Where where;
  if (device_type == kDLROCM) {
    where = dl::Where::Device;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipMalloc(&data, bytes));
  } else if (device_type == kDLCPU) {
    where = dl::Where::Host;
    data = std::malloc(bytes);
    if (!data) return fail(CST_ERR_HIP, "host allocation failed");
  } else {
    return fail(CST_ERR_ARG, "export device type must be kDLCPU (1) or kDLROCM (10)");
  }
  *out = dl::make_managed(data, where, device, dt, shape, ndim, release_export);
  if (!*out) return fail(CST_ERR_HIP, "host allocation failed");
  return CST_OK;
}

}  // namespace

extern "C" {

int cst_member_result_export(cst_session* s, int slot, int hi, double scale, int device_type,
                             DLManagedTensor** phasor, DLManagedTensor** pmax) {
  if (!phasor) return fail(CST_ERR_ARG, "null output");
  *phasor = nullptr;
  if (pmax) *pmax = nullptr;
  if (int r = check_slot(s, slot)) return r;
  const int* rb = s->members[slot].dev.rec;
  const int64_t shape[3] = {rb[1] - rb[0], rb[3] - rb[2], rb[5] - rb[4]};
  DLManagedTensor *a = nullptr, *b = nullptr;
  if (int r = alloc_export(device_type, s->device, dl::kC64, shape, 3, &a)) return r;
  if (pmax) {
    if (int r = alloc_export(device_type, s->device, dl::kF32, shape, 3, &b)) {
      a->deleter(a);
      return r;
    }
  }
  if (int r = cst_member_result_into(s, slot, hi, scale, &a->dl_tensor, b ? &b->dl_tensor : nullptr)) {
    a->deleter(a);
    if (b) b->deleter(b);
    return r;
  }
  *phasor = a;
  if (pmax) *pmax = b;
  return CST_OK;
}

int cst_dl_alloc(int device_type, int device, int dtype_code, int dtype_bits, const int64_t* shape,
                 int ndim, DLManagedTensor** out) {
  if (!out || (ndim > 0 && !shape) || ndim < 0 || ndim > 8) return fail(CST_ERR_ARG, "invalid arguments");
  *out = nullptr;
  for (int i = 0; i < ndim; ++i)
    if (shape[i] < 0) return fail(CST_ERR_ARG, "negative extent");
  if (dtype_bits % 8 != 0 || dtype_bits < 8 || dtype_bits > 128) return fail(CST_ERR_ARG, "invalid dtype bits");
  const DLDataType dt{(uint8_t)dtype_code, (uint8_t)dtype_bits, 1};
  return alloc_export(device_type, device, dt, shape, ndim, out);
}

int cst_dl_copy(const DLTensor* src, DLTensor* dst) {
  if (!src || !dst) return fail(CST_ERR_ARG, "null tensor");
  int64_t n = 1;
  for (int i = 0; i < src->ndim; ++i) n *= src->shape ? src->shape[i] : 0;
  const int dev = src->device.device_type == kDLROCM ? src->device.device_id
                  : dst->device.device_type == kDLROCM ? dst->device.device_id : 0;
  dl::View a, b;
  if (int r = check_dl(src, "src", src->dtype, n, dev, &a)) return r;
  if (int r = check_dl(dst, "dst", src->dtype, n, dev, &b)) return r;
  const size_t bytes = (size_t)n * (src->dtype.bits / 8) * src->dtype.lanes;
  if (bytes == 0) return CST_OK;
  if (a.where == dl::Where::Host && b.where == dl::Where::Host) {
    std::memmove(b.data, a.data, bytes);
    return CST_OK;
  }
  HIPCHK(hipSetDevice(dev));
  HIPCHK(hipMemcpy(b.data, a.data, bytes, hipMemcpyDefault));
  return CST_OK;
}

int cst_member_get_state_into(cst_session* s, int slot, DLTensor* p, DLTensor* u0, DLTensor* u1,
                              DLTensor* u2, int64_t* n) {
  if (int r = check_slot(s, slot)) return r;
  Member& m = s->members[slot];
  const int64_t P = (int64_t)s->s.nx * s->s.ny * s->s.nz;
  DLTensor* t[4] = {p, u0, u1, u2};
  const char* names[4] = {"p", "u0", "u1", "u2"};
  dl::View v[4];
  for (int i = 0; i < 4; ++i)
    if (t[i])
      if (int r = check_dl(t[i], names[i], dl::kF32, P, s->device, &v[i])) return r;
  HIPCHK(hipSetDevice(s->device));
  if (p) HIPCHK(hipMemcpyAsync(v[0].data, m.dev.p, P * 4, from_device(v[0]), s->stream));
  // velocity: C2R of the state spectrum into the (scratch) gradient spectrum's memory
  StepArgs a = single_member(s, slot);
  for (int ax = 0; ax < 3; ++ax) {
    if (!t[ax + 1]) continue;
    if (!launch_zconv(s->s.nz / 2, a, s->tw_z, s->rt, ax, -1, s->stream))
      return fail(CST_ERR_UNSUPPORTED, "no kernel for nz");
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(v[ax + 1].data, m.dev.sg[ax], P * 4, from_device(v[ax + 1]), s->stream));
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  if (n) *n = m.n;
  return CST_OK;
}

int cst_member_get_state(cst_session* s, int slot, float* p, float* u0, float* u1, float* u2,
                         int64_t* n) {
  if (!s) return fail(CST_ERR_ARG, "null session");
  int64_t shp[3] = {s->s.nx, s->s.ny, s->s.nz};
  float* src[4] = {p, u0, u1, u2};
  DLTensor t[4];
  for (int i = 0; i < 4; ++i) t[i] = dl::host_tensor(src[i], dl::kF32, shp, 3);
  return cst_member_get_state_into(s, slot, p ? &t[0] : nullptr, u0 ? &t[1] : nullptr,
                                   u1 ? &t[2] : nullptr, u2 ? &t[3] : nullptr, n);
}

int cst_member_set_state_dl(cst_session* s, int slot, int64_t n, const DLTensor* p, const DLTensor* u0,
                            const DLTensor* u1, const DLTensor* u2) {
  if (int r = check_slot(s, slot)) return r;
  if (n < 0) return fail(CST_ERR_ARG, "negative step");
  Member& m = s->members[slot];
  const int64_t P = (int64_t)s->s.nx * s->s.ny * s->s.nz;
  const DLTensor* t[4] = {p, u0, u1, u2};
  const char* names[4] = {"p", "u0", "u1", "u2"};
  dl::View v[4];
  for (int i = 0; i < 4; ++i)
    if (int r = check_dl(t[i], names[i], dl::kF32, P, s->device, &v[i])) return r;
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipMemcpyAsync(m.dev.p, v[0].data, P * 4, to_device(v[0]), s->stream));
  StepArgs a = single_member(s, slot);
  for (int ax = 0; ax < 3; ++ax) {
    HIPCHK(hipMemcpyAsync(m.dev.sg[ax], v[ax + 1].data, P * 4, to_device(v[ax + 1]), s->stream));
    if (!launch_zconv(s->s.nz / 2, a, s->tw_z, s->rt, ax, +1, s->stream))
      return fail(CST_ERR_UNSUPPORTED, "no kernel for nz");
    HIPCHK(hipGetLastError());
  }
  if (int r = restart_record(s, m, n)) return r;
  HIPCHK(hipStreamSynchronize(s->stream));
  m.n = n;
  m.primed = false;
  return CST_OK;
}

int cst_member_set_state(cst_session* s, int slot, int64_t n, const float* p, const float* u0,
                         const float* u1, const float* u2) {
  if (!s) return fail(CST_ERR_ARG, "null session");
  if (!p || !u0 || !u1 || !u2) return fail(CST_ERR_ARG, "null state array or negative step");
  int64_t shp[3] = {s->s.nx, s->s.ny, s->s.nz};
  const float* src[4] = {p, u0, u1, u2};
  DLTensor t[4];
  for (int i = 0; i < 4; ++i) t[i] = dl::host_tensor(const_cast<float*>(src[i]), dl::kF32, shp, 3);
  return cst_member_set_state_dl(s, slot, n, &t[0], &t[1], &t[2], &t[3]);
}

int cst_member_get_density_into(cst_session* s, int slot, DLTensor* rho0, DLTensor* rho1, DLTensor* rho2) {
  if (int r = check_slot(s, slot)) return r;
  if (s->scheme != CST_SCHEME_STAGGERED) return fail(CST_ERR_STATE, "not a staggered session");
  Member& m = s->members[slot];
  const int64_t P = (int64_t)s->s.nx * s->s.ny * s->s.nz;
  DLTensor* t[3] = {rho0, rho1, rho2};
  const char* names[3] = {"rho0", "rho1", "rho2"};
  dl::View v[3];
  for (int i = 0; i < 3; ++i)
    if (t[i])
      if (int r = check_dl(t[i], names[i], dl::kF32, P, s->device, &v[i])) return r;
  HIPCHK(hipSetDevice(s->device));
  for (int i = 0; i < 3; ++i)
    if (t[i]) HIPCHK(hipMemcpyAsync(v[i].data, m.dev.stg.rho[i], P * 4, from_device(v[i]), s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return CST_OK;
}

int cst_member_set_density_dl(cst_session* s, int slot, const DLTensor* rho0, const DLTensor* rho1,
                              const DLTensor* rho2) {
  if (int r = check_slot(s, slot)) return r;
  if (s->scheme != CST_SCHEME_STAGGERED) return fail(CST_ERR_STATE, "not a staggered session");
  Member& m = s->members[slot];
  const int64_t P = (int64_t)s->s.nx * s->s.ny * s->s.nz;
  const DLTensor* t[3] = {rho0, rho1, rho2};
  const char* names[3] = {"rho0", "rho1", "rho2"};
  dl::View v[3];
  for (int i = 0; i < 3; ++i)
    if (int r = check_dl(t[i], names[i], dl::kF32, P, s->device, &v[i])) return r;
  HIPCHK(hipSetDevice(s->device));
  for (int i = 0; i < 3; ++i)
    HIPCHK(hipMemcpyAsync(m.dev.stg.rho[i], v[i].data, P * 4, to_device(v[i]), s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return CST_OK;
}

static size_t raw_state_bytes(const cst_session* s) {
  const Shape& sh = s->s;
  const size_t pb = (size_t)sh.nx * sh.ny * sh.nz * 4;
  return pb + 3 * (size_t)sh.hs * 8 + (s->scheme == CST_SCHEME_STAGGERED ? 3 * pb : 0);
}

int cst_member_state_size(cst_session* s, size_t* bytes) {
  if (!s || !bytes) return fail(CST_ERR_ARG, "null argument");
  *bytes = raw_state_bytes(s);
  return CST_OK;
}

int cst_member_get_raw(cst_session* s, int slot, void* buf, size_t bytes, int64_t* n) {
  if (int r = check_slot(s, slot)) return r;
  if (!buf || bytes != raw_state_bytes(s)) return fail(CST_ERR_ARG, "buffer size differs from cst_member_state_size");
  Member& m = s->members[slot];
  const size_t pb = (size_t)s->s.nx * s->s.ny * s->s.nz * 4, hb = (size_t)s->s.hs * 8;
  HIPCHK(hipSetDevice(s->device));
  char* dst = static_cast<char*>(buf);
  HIPCHK(hipMemcpyAsync(dst, m.dev.p, pb, hipMemcpyDeviceToHost, s->stream));
  for (int ax = 0; ax < 3; ++ax)
    HIPCHK(hipMemcpyAsync(dst + pb + ax * hb, m.dev.su[ax], hb, hipMemcpyDeviceToHost, s->stream));
  if (s->scheme == CST_SCHEME_STAGGERED)
    for (int ax = 0; ax < 3; ++ax)
      HIPCHK(hipMemcpyAsync(dst + pb + 3 * hb + ax * pb, m.dev.stg.rho[ax], pb, hipMemcpyDeviceToHost,
                            s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (n) *n = m.n;
  return CST_OK;
}

int cst_member_set_raw(cst_session* s, int slot, const void* buf, size_t bytes, int64_t n) {
  if (int r = check_slot(s, slot)) return r;
  if (!buf || bytes != raw_state_bytes(s)) return fail(CST_ERR_ARG, "buffer size differs from cst_member_state_size");
  if (n < 0) return fail(CST_ERR_ARG, "negative step");
  Member& m = s->members[slot];
  const size_t pb = (size_t)s->s.nx * s->s.ny * s->s.nz * 4, hb = (size_t)s->s.hs * 8;
  HIPCHK(hipSetDevice(s->device));
  const char* src = static_cast<const char*>(buf);
  HIPCHK(hipMemcpyAsync(m.dev.p, src, pb, hipMemcpyHostToDevice, s->stream));
  for (int ax = 0; ax < 3; ++ax)
    HIPCHK(hipMemcpyAsync(m.dev.su[ax], src + pb + ax * hb, hb, hipMemcpyHostToDevice, s->stream));
  if (s->scheme == CST_SCHEME_STAGGERED)
    for (int ax = 0; ax < 3; ++ax)
      HIPCHK(hipMemcpyAsync(m.dev.stg.rho[ax], src + pb + 3 * hb + ax * pb, pb, hipMemcpyHostToDevice,
                            s->stream));
  if (int r = restart_record(s, m, n)) return r;
  HIPCHK(hipStreamSynchronize(s->stream));
  m.n = n;
  m.primed = false;
  return CST_OK;
}

int cst_session_set_variant(cst_session* s, int pass, int variant) {
  if (!s) return fail(CST_ERR_ARG, "null session");
  bool ok = false;
  if (pass == 0 || pass == 1) ok = variant >= 0 && variant <= 2;
  if (pass == 2) ok = (variant >= 0 && variant <= 2) || (variant >= 4 && variant <= 6);
  if (pass == 3) ok = variant == 0 || variant == 1 || variant == 4 || variant == 5;
  if (!ok)
    return fail(CST_ERR_ARG, "pass 0 (y) / 1 (x): tile variant 0..2; pass 2 (zu): layout 0..2 + 4 x radix "
                             "plan 0..1; pass 3 (zp): layout 0..1 + 4 x radix plan 0..1");
  s->variant[pass] = variant;  // cached graphs are keyed by the variants they captured
  return CST_OK;
}

int cst_session_set_graphs(cst_session* s, int on) {
  if (!s) return fail(CST_ERR_ARG, "null session");
  HIPCHK(hipSetDevice(s->device));
  destroy_graphs(s);
  s->graphs = on != 0;
  return CST_OK;
}

int cst_session_graph_stats(cst_session* s, int* enabled, int64_t* launches, int64_t* captures) {
  if (!s || !enabled || !launches || !captures) return fail(CST_ERR_ARG, "null argument");
  *enabled = s->graphs ? 1 : 0;
  *launches = s->graph_launches;
  *captures = s->graph_captures;
  return CST_OK;
}

int cst_session_set_profiling(cst_session* s, int on) {
  if (!s) return fail(CST_ERR_ARG, "null session");
  HIPCHK(hipSetDevice(s->device));
  return s->prof.enable(on != 0, s->stream);
}

int cst_session_kernel_time(cst_session* s, int kind, double* total_ms, int64_t* launches) {
  if (!s || !total_ms || !launches) return fail(CST_ERR_ARG, "null argument");
  if (kind < 0 || kind >= Profiler::kKinds) return fail(CST_ERR_ARG, "unknown kernel kind");
  HIPCHK(hipSetDevice(s->device));
  if (int r = s->prof.resolve()) return r;
  *total_ms = s->prof.ms[kind];
  *launches = s->prof.count[kind];
  return CST_OK;
}

int cst_thermal_heat_from_member(cst_thermal* t, cst_session* s, int slot, int n_terms,
                                 const int* harmonic_index, const double* factor, const float* w,
                                 double scale, float q_offset) {
  if (!t || !harmonic_index || !factor || !w) return fail(CST_ERR_ARG, "null argument");
  if (int r = check_slot(s, slot)) return r;
  if (!t->ready) return fail(CST_ERR_STATE, "thermal session not set up");
  if (t->device != s->device) return fail(CST_ERR_ARG, "thermal and acoustic sessions on different devices");
  Member& m = s->members[slot];
  if (n_terms < 1 || n_terms > m.nh) return fail(CST_ERR_ARG, "n_terms must be 1..n_harmonics");
  const int* rc = m.dev.rec;
  if (rc[1] - rc[0] != t->nx || rc[3] - rc[2] != t->ny || rc[5] - rc[4] != t->nz)
    return fail(CST_ERR_ARG, "thermal grid does not match the member's record region");
  HeatJoin j{};
  for (int k = 0; k < n_terms; ++k) {
    const int hi = harmonic_index[k];
    if (hi < 0 || hi >= m.nh) return fail(CST_ERR_ARG, "harmonic index out of range");
    if (!m.finalized[hi]) return fail(CST_ERR_STATE, "phasor not finalized: call cst_member_result first");
    j.p[k] = m.dev.comp[hi];
    j.f[k] = factor[k];
  }
  j.nt = n_terms;
  j.n = m.rec_count;
  j.perf = t->perf;
  j.qoff = q_offset;
  j.scale = scale;
  j.q = t->q;
  HIPCHK(hipSetDevice(s->device));
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipStreamSynchronize(s->io));  // the finalized phasors are complete
  // w is staged in the thermal theta[1] buffer (free until the first thermal sub-step)
  float* wd = t->theta[1 - t->cur];
  HIPCHK(hipMemcpyAsync(wd, w, (size_t)j.n * 4, hipMemcpyHostToDevice, t->stream));
  j.w = wd;
  const unsigned blocks = (unsigned)((j.n + 255) / 256);
  hipLaunchKernelGGL(k_heat_join, dim3(blocks), dim3(256), 0, t->stream, j);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(t->stream));
  t->have_q = true;
  return CST_OK;
}

}  // extern "C"
