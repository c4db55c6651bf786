// Error reporting, device-memory carving and kernel timing shared by the host code.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/caustica_hip.h"

namespace cst {

inline std::string& last_error() {
  static thread_local std::string msg;
  return msg;
}

inline int fail(int code, const std::string& msg) {
  last_error() = msg;
  return code;
}

inline int copy_last_error(char* buf, size_t n) {
  if (!buf || n == 0) return CST_ERR_ARG;
  const std::string& m = last_error();
  const size_t k = m.size() < n - 1 ? m.size() : n - 1;
  std::memcpy(buf, m.data(), k);
  buf[k] = '\0';
  return CST_OK;
}

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return ::cst::fail(CST_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Offsets of 256-byte aligned pieces inside one allocation.
struct Carver {
  size_t total = 0;
  size_t take(int64_t bytes) {
    const size_t off = total;
    total += ((size_t)(bytes > 0 ? bytes : 0) + 255) & ~(size_t)255;
    return off;
  }
};

// Per-kind kernel durations from HIP events recorded around each launch.
struct Profiler {
  static constexpr int kKinds = 8;
  bool on = false;
  double ms[kKinds] = {};
  int64_t count[kKinds] = {};
  struct Rec {
    int kind;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;

  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }

  template <class F>
  bool run(int kind, hipStream_t st, F f) {
    if (!on) return f();
    hipEvent_t a = get(), b = get();
    (void)hipEventRecord(a, st);
    const bool ok = f();
    (void)hipEventRecord(b, st);
    pending.push_back({kind, a, b});
    if (pending.size() >= 8192) resolve();
    return ok;
  }

  int resolve() {
    for (auto& r : pending) {
      HIPCHK(hipEventSynchronize(r.b));
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
      ms[r.kind] += t;
      count[r.kind] += 1;
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pending.clear();
    return CST_OK;
  }

  int enable(bool v, hipStream_t) {
    if (int r = resolve()) return r;
    on = v;
    for (int i = 0; i < kKinds; ++i) {
      ms[i] = 0.0;
      count[i] = 0;
    }
    return CST_OK;
  }

  void release() {
    resolve();
    for (auto e : pool) (void)hipEventDestroy(e);
    pool.clear();
  }
};

}  // namespace cst
