// Argument and ownership layer of the DLPack entry points: checks that a borrowed tensor
// is what an entry point needs (dtype, element count or shape, compact C order, a device
// the call may read) and builds the managed tensors the library exports. Plain C++ with
// no HIP dependency, so it builds with the host compiler alone (and under the address
// sanitizer for its tests); where memory lives and how it is released is the caller's
// business, passed in as a release function.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/cst_dlpack.h"

namespace cst::dl {

enum class Where { Host = 0, Device = 1 };

// A checked, borrowed tensor: its first element and where it lives.
struct View {
  void* data = nullptr;
  Where where = Where::Host;
  int device = 0;
  int64_t count = 0;
};

constexpr DLDataType kF32{kDLFloat, 32, 1};
constexpr DLDataType kF64{kDLFloat, 64, 1};
constexpr DLDataType kC64{kDLComplex, 64, 1};
constexpr DLDataType kI32{kDLInt, 32, 1};

// Accepts host memory (kDLCPU, kDLCUDAHost, kDLROCMHost) anywhere and device memory
// (kDLROCM) on `device` only. `shape` (ndim entries) pins the exact shape when given;
// otherwise only the element count `count` must match (any rank). Strides must be NULL
// or describe compact C order; size-1 dimensions may carry any stride. Returns false
// with a message naming `what` on any mismatch.
bool check(const DLTensor* t, const char* what, DLDataType want, int64_t count, const int64_t* shape,
           int ndim, int device, View* out, std::string* err);

// Releases an exported tensor's data; `where` and `device` are what make_managed got.
using ReleaseFn = void (*)(void* data, Where where, int device);

// A managed tensor over `data` (C order, no strides) whose deleter calls release(data)
// and frees the tensor and its shape; nullptr if the host allocation fails (data is then
// released at once).
DLManagedTensor* make_managed(void* data, Where where, int device, DLDataType dtype,
                              const int64_t* shape, int ndim, ReleaseFn release);

// A plain host tensor view (kDLCPU, compact C order) over caller memory: what the
// plain-pointer entry points wrap their arguments in. `shape` must outlive the use.
DLTensor host_tensor(void* data, DLDataType dtype, int64_t* shape, int ndim);

}  // namespace cst::dl
