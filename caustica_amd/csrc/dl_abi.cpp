// Argument and ownership layer of the DLPack entry points (dl_abi.h).
#include "dl_abi.h"

#include <cstdlib>
#include <cstring>
#include <new>

namespace cst::dl {

namespace {

bool same(DLDataType a, DLDataType b) { return a.code == b.code && a.bits == b.bits && a.lanes == b.lanes; }

std::string dtype_name(DLDataType t) {
  const char* c = t.code == kDLFloat ? "float" : t.code == kDLComplex ? "complex"
                 : t.code == kDLInt ? "int" : t.code == kDLUInt ? "uint" : "code?";
  std::string s = std::string(c) + std::to_string((int)t.bits);
  if (t.lanes != 1) s += "x" + std::to_string((int)t.lanes);
  return s;
}

struct Ctx {
  int64_t shape[8];
  ReleaseFn release;
  Where where;
  int device;
};

void managed_deleter(DLManagedTensor* self) {
  if (!self) return;
  Ctx* c = static_cast<Ctx*>(self->manager_ctx);
  if (c) {
    if (c->release && self->dl_tensor.data) c->release(self->dl_tensor.data, c->where, c->device);
    delete c;
  }
  delete self;
}

}  // namespace

bool check(const DLTensor* t, const char* what, DLDataType want, int64_t count, const int64_t* shape,
           int ndim, int device, View* out, std::string* err) {
  auto bad = [&](const std::string& m) {
    if (err) *err = std::string(what) + ": " + m;
    return false;
  };
  if (!t) return bad("null tensor");
  if (!t->data) return bad("null data pointer");
  if (!same(t->dtype, want)) return bad("dtype " + dtype_name(t->dtype) + ", expected " + dtype_name(want));
  if (t->ndim < 0 || t->ndim > 8) return bad("rank " + std::to_string(t->ndim) + " outside 0..8");
  if (t->ndim > 0 && !t->shape) return bad("null shape");
  Where where;
  switch (t->device.device_type) {
    case kDLCPU:
    case kDLCUDAHost:
    case kDLROCMHost:
      where = Where::Host;
      break;
    case kDLROCM:
      where = Where::Device;
      if (t->device.device_id != device)
        return bad("device tensor on GPU " + std::to_string(t->device.device_id) + ", the call runs on GPU " +
                   std::to_string(device));
      break;
    default:
      return bad("device type " + std::to_string(t->device.device_type) + " is neither host memory nor a ROCm GPU");
  }
  int64_t n = 1;
  for (int i = 0; i < t->ndim; ++i) {
    if (t->shape[i] < 0) return bad("negative extent");
    n *= t->shape[i];
  }
  if (shape) {
    if (t->ndim != ndim) return bad("rank " + std::to_string(t->ndim) + ", expected " + std::to_string(ndim));
    for (int i = 0; i < ndim; ++i)
      if (t->shape[i] != shape[i])
        return bad("extent " + std::to_string(t->shape[i]) + " on axis " + std::to_string(i) + ", expected " +
                   std::to_string(shape[i]));
  } else if (n != count) {
    return bad(std::to_string(n) + " elements, expected " + std::to_string(count));
  }
  if (t->strides) {
    int64_t expect = 1;
    for (int i = t->ndim - 1; i >= 0; --i) {
      if (t->shape[i] != 1 && t->strides[i] != expect) return bad("not compact C order (strided view)");
      expect *= t->shape[i];
    }
  }
  const size_t esize = (size_t)t->dtype.bits / 8;
  if (t->byte_offset % esize != 0) return bad("byte offset not a multiple of the element size");
  out->data = static_cast<char*>(t->data) + t->byte_offset;
  out->where = where;
  out->device = where == Where::Device ? t->device.device_id : 0;
  out->count = n;
  return true;
}

DLManagedTensor* make_managed(void* data, Where where, int device, DLDataType dtype, const int64_t* shape,
                              int ndim, ReleaseFn release) {
  if (ndim < 0 || ndim > 8) {
    if (release && data) release(data, where, device);
    return nullptr;
  }
  Ctx* c = new (std::nothrow) Ctx();
  DLManagedTensor* m = c ? new (std::nothrow) DLManagedTensor() : nullptr;
  if (!c || !m) {
    delete c;
    if (release && data) release(data, where, device);
    return nullptr;
  }
  for (int i = 0; i < ndim; ++i) c->shape[i] = shape[i];
  c->release = release;
  c->where = where;
  c->device = device;
  m->dl_tensor.data = data;
  m->dl_tensor.device.device_type = where == Where::Device ? kDLROCM : kDLCPU;
  m->dl_tensor.device.device_id = where == Where::Device ? device : 0;
  m->dl_tensor.ndim = ndim;
  m->dl_tensor.dtype = dtype;
  m->dl_tensor.shape = c->shape;
  m->dl_tensor.strides = nullptr;
  m->dl_tensor.byte_offset = 0;
  m->manager_ctx = c;
  m->deleter = managed_deleter;
  return m;
}

DLTensor host_tensor(void* data, DLDataType dtype, int64_t* shape, int ndim) {
  DLTensor t{};
  t.data = data;
  t.device.device_type = kDLCPU;
  t.device.device_id = 0;
  t.ndim = ndim;
  t.dtype = dtype;
  t.shape = shape;
  t.strides = nullptr;
  t.byte_offset = 0;
  return t;
}

}  // namespace cst::dl
