// pmc_calib.hip -- TEST INFRASTRUCTURE: calibration of rocprofv3's HBM byte
// counters (never part of the product).  5 hipMemcpyAsync device-to-device
// copies of 512 MiB (the runtime's own copy kernel), then 5 launches of a
// plain vector copy kernel over the same 512 MiB, so that FETCH_SIZE /
// WRITE_SIZE per dispatch can be set against the known 512 MiB read and
// 512 MiB written by each.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                        \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      std::fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
      std::exit(1);                                                                  \
    }                                                                                \
  } while (0)

typedef unsigned u4 __attribute__((ext_vector_type(4)));

__global__ void k_plain_copy(const u4 *__restrict__ a, u4 *__restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

int main() {
  const size_t bytes = (size_t)512 << 20;
  void *a, *b;
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  CK(hipMemset(a, 1, bytes));
  CK(hipMemset(b, 2, bytes));
  CK(hipDeviceSynchronize());
  for (int r = 0; r < 5; ++r) CK(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, 0));
  CK(hipDeviceSynchronize());
  const size_t n = bytes / 16;
  for (int r = 0; r < 5; ++r) {
    hipLaunchKernelGGL(k_plain_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0,
                       (const u4 *)a, (u4 *)b, n);
    CK(hipGetLastError());
  }
  CK(hipDeviceSynchronize());
  std::printf("calib done: 5 x hipMemcpyAsync D2D + 5 x k_plain_copy of %zu bytes\n", bytes);
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
