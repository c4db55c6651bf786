/* cst_dlpack.h -- the DLPack tensor ABI (DLPack 0.8 layout) as the caustica_hip library
 * takes and returns it.
 *
 * DLPack is the cross-framework in-memory tensor description (numpy, PyTorch and
 * others export and import it). These are its C structures, field for field; when the
 * project's own dlpack.h is included first (it defines DLPACK_VERSION), that header's
 * definitions are used instead.
 */
#ifndef CST_DLPACK_H
#define CST_DLPACK_H

#include <stdint.h>

#ifndef DLPACK_VERSION

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  kDLCPU = 1,
  kDLCUDA = 2,
  kDLCUDAHost = 3,
  kDLROCM = 10,
  kDLROCMHost = 11,
} DLDeviceType;

typedef struct {
  int32_t device_type; /* DLDeviceType */
  int32_t device_id;
} DLDevice;

typedef enum {
  kDLInt = 0,
  kDLUInt = 1,
  kDLFloat = 2,
  kDLBfloat = 4,
  kDLComplex = 5,
  kDLBool = 6,
} DLDataTypeCode;

typedef struct {
  uint8_t code;  /* DLDataTypeCode */
  uint8_t bits;
  uint16_t lanes;
} DLDataType;

typedef struct {
  void* data;
  DLDevice device;
  int32_t ndim;
  DLDataType dtype;
  int64_t* shape;
  int64_t* strides; /* in elements; NULL: compact C order */
  uint64_t byte_offset;
} DLTensor;

typedef struct DLManagedTensor {
  DLTensor dl_tensor;
  void* manager_ctx;
  void (*deleter)(struct DLManagedTensor* self);
} DLManagedTensor;

#ifdef __cplusplus
}
#endif

#endif /* DLPACK_VERSION */
#endif /* CST_DLPACK_H */
