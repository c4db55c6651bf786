// Template instantiation + launch of the k-space pass kernels: supported lengths.
#include "kspace_kernels.h"
#include "kspace_launch.h"

namespace cst {

bool length_supported(int L) {
  switch (L) {
#define CST_SUP(N) case N:
    CST_LENGTHS(CST_SUP)
    return true;
#undef CST_SUP
    default:
      return false;
  }
}

}  // namespace cst
