// batch_kernels_c64.hip -- every batch kernel instance for complex<float>.
#include "batch_kernels.h"

namespace hpttx {

void batch_kernels_c64(const void **tab) {
  bk::fill_kernels<Elem<float, 2>, float, 2, 2, true, 0>(tab);
}

}  // namespace hpttx
