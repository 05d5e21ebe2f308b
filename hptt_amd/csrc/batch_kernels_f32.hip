// batch_kernels_f32.hip -- every batch kernel instance for float.
#include "batch_kernels.h"

namespace hpttx {

void batch_kernels_f32(const void **tab) {
  bk::fill_kernels<Elem<float, 1>, float, 1, 4, false, 0>(tab);
}

}  // namespace hpttx
