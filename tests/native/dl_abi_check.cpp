// Checks of the DLPack argument/ownership layer (caustica_amd/csrc/dl_abi.cpp), built
// with the host compiler under the address and undefined-behaviour sanitizers by
// tests/test_dl_abi_asan.py: every accepted and refused tensor form, and the export
// lifecycle (the deleter releases the data exactly once and frees the tensor).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../caustica_amd/csrc/dl_abi.h"

using namespace cst::dl;

static int failures = 0;
#define EXPECT(c)                                                   \
  do {                                                              \
    if (!(c)) {                                                     \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
      ++failures;                                                   \
    }                                                               \
  } while (0)

static int released = 0;
static void release_host(void* data, Where, int) {
  ++released;
  std::free(data);
}

int main() {
  float buf[24] = {};
  int64_t shape[3] = {2, 3, 4};
  View v;
  std::string err;

  // accepted: compact host tensor, count match, any rank with the right element count
  DLTensor t = host_tensor(buf, kF32, shape, 3);
  EXPECT(check(&t, "a", kF32, 24, nullptr, 0, 0, &v, &err));
  EXPECT(v.data == buf && v.where == Where::Host && v.count == 24);
  int64_t flat[1] = {24};
  DLTensor f = host_tensor(buf, kF32, flat, 1);
  EXPECT(check(&f, "a", kF32, 24, nullptr, 0, 0, &v, &err));
  EXPECT(check(&t, "a", kF32, 0, shape, 3, 0, &v, &err));  // exact shape
  // explicit compact strides, and size-1 axes with any stride
  int64_t st[3] = {12, 4, 1};
  t.strides = st;
  EXPECT(check(&t, "a", kF32, 24, nullptr, 0, 0, &v, &err));
  int64_t shape1[3] = {1, 6, 4}, st1[3] = {999, 4, 1};
  DLTensor o = host_tensor(buf, kF32, shape1, 3);
  o.strides = st1;
  EXPECT(check(&o, "a", kF32, 24, nullptr, 0, 0, &v, &err));
  // byte offset moves the first element
  DLTensor b = host_tensor(buf, kF32, flat, 1);
  int64_t flat23[1] = {23};
  b.shape = flat23;
  b.byte_offset = 4;
  EXPECT(check(&b, "a", kF32, 23, nullptr, 0, 0, &v, &err) && v.data == buf + 1);
  // device tensor on the call's GPU is accepted, another GPU refused
  DLTensor d = host_tensor(buf, kF32, shape, 3);
  d.device.device_type = kDLROCM;
  d.device.device_id = 1;
  EXPECT(check(&d, "a", kF32, 24, nullptr, 0, 1, &v, &err) && v.where == Where::Device && v.device == 1);
  EXPECT(!check(&d, "a", kF32, 24, nullptr, 0, 0, &v, &err) && err.find("GPU 1") != std::string::npos);

  // refused: null tensor / data / shape, dtype, count, rank, strided view, device type
  EXPECT(!check(nullptr, "a", kF32, 24, nullptr, 0, 0, &v, &err));
  DLTensor z = host_tensor(nullptr, kF32, shape, 3);
  EXPECT(!check(&z, "a", kF32, 24, nullptr, 0, 0, &v, &err) && err.find("null data") != std::string::npos);
  DLTensor ns = host_tensor(buf, kF32, nullptr, 3);
  EXPECT(!check(&ns, "a", kF32, 24, nullptr, 0, 0, &v, &err));
  DLTensor w = host_tensor(buf, kF64, shape, 3);
  EXPECT(!check(&w, "maps", kF32, 24, nullptr, 0, 0, &v, &err) && err.find("float64") != std::string::npos &&
         err.rfind("maps", 0) == 0);
  DLTensor c = host_tensor(buf, kF32, shape, 3);
  EXPECT(!check(&c, "a", kF32, 25, nullptr, 0, 0, &v, &err) && err.find("24 elements") != std::string::npos);
  int64_t other[3] = {3, 2, 4};
  EXPECT(!check(&c, "a", kF32, 0, other, 3, 0, &v, &err));
  EXPECT(!check(&c, "a", kF32, 0, shape, 2, 0, &v, &err));
  int64_t sv[3] = {12, 8, 2};  // every second element
  c.strides = sv;
  EXPECT(!check(&c, "a", kF32, 24, nullptr, 0, 0, &v, &err) && err.find("strided") != std::string::npos);
  DLTensor u = host_tensor(buf, kF32, shape, 3);
  u.device.device_type = kDLCUDA;
  EXPECT(!check(&u, "a", kF32, 24, nullptr, 0, 0, &v, &err));
  DLTensor bo = host_tensor(buf, kF32, flat23, 1);
  bo.byte_offset = 2;
  EXPECT(!check(&bo, "a", kF32, 23, nullptr, 0, 0, &v, &err));
  int64_t neg[1] = {-1};
  DLTensor ng = host_tensor(buf, kF32, neg, 1);
  EXPECT(!check(&ng, "a", kF32, -1, nullptr, 0, 0, &v, &err));

  // export lifecycle: the deleter releases the data once and frees the tensor
  for (int rep = 0; rep < 3; ++rep) {
    void* data = std::malloc(24 * 8);
    const int64_t es[3] = {2, 3, 4};
    DLManagedTensor* m = make_managed(data, Where::Host, 0, kC64, es, 3, release_host);
    EXPECT(m != nullptr);
    EXPECT(m->dl_tensor.data == data && m->dl_tensor.ndim == 3 && m->dl_tensor.device.device_type == kDLCPU);
    EXPECT(m->dl_tensor.shape[0] == 2 && m->dl_tensor.shape[2] == 4 && m->dl_tensor.strides == nullptr);
    std::memset(data, 0, 24 * 8);
    m->deleter(m);
  }
  EXPECT(released == 3);
  // a device export records its GPU; a rank above 8 is refused and releases the data
  DLManagedTensor* md = make_managed(std::malloc(16), Where::Device, 3, kF32, shape, 3, release_host);
  EXPECT(md && md->dl_tensor.device.device_type == kDLROCM && md->dl_tensor.device.device_id == 3);
  md->deleter(md);
  EXPECT(make_managed(std::malloc(16), Where::Host, 0, kF32, shape, 9, release_host) == nullptr);
  EXPECT(released == 5);

  if (failures) return 1;
  std::printf("dl_abi ok\n");
  return 0;
}
