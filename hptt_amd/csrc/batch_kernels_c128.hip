// batch_kernels_c128.hip -- every batch kernel instance for complex<double>.
#include "batch_kernels.h"

namespace hpttx {

void batch_kernels_c128(const void **tab) {
  bk::fill_kernels<Elem<double, 2>, double, 2, 1, true, 0>(tab);
}

}  // namespace hpttx
