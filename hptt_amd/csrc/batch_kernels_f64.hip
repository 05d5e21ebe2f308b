// batch_kernels_f64.hip -- every batch kernel instance for double.
#include "batch_kernels.h"

namespace hpttx {

void batch_kernels_f64(const void **tab) {
  bk::fill_kernels<Elem<double, 1>, double, 1, 2, false, 0>(tab);
}

}  // namespace hpttx
