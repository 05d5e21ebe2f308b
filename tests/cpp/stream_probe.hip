// stream_probe.hip -- TEST INFRASTRUCTURE: the HBM streaming ceiling of the
// GPU it runs on, the reference point for the transposition kernels'
// bandwidth (never part of the product).
//
// Kernels: read-only, write-only, copy (1 read : 1 write) and axpby (2 reads
// : 1 write, the byte mix of a transposition with beta), over two buffers of
// the size given in MiB (default 2048).  Variants: U 16-byte vectors per
// thread, block size, one-shot grid (one block per chunk of U * block
// vectors) or a persistent grid-stride loop, non-temporal or default loads
// and stores.  Also hipMemcpyAsync device to device.  Every store is
// guarded by idx < n (n = number of 16-byte vectors); loads are guarded the
// same way (bf=0), or issued for every slot with out-of-range slots reading
// vector 0 (bf=1: no branch around a load, so all U loads of a thread are in
// flight together; under a branch per load the compiler waits for each load
// before the next).  Prints GB/s per variant: bytes moved / mean event-timed
// duration of 20 back-to-back launches after one warm-up launch.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

typedef float f4 __attribute__((ext_vector_type(4)));

#define CK(x)                                                                  \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      std::fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
      std::exit(1);                                                            \
    }                                                                          \
  } while (0)

template <bool NT>
__device__ __forceinline__ f4 ld(const f4 *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <bool NT>
__device__ __forceinline__ void st(f4 *p, f4 v) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// MODE: 0 read, 1 write, 2 copy, 3 axpby
template <int MODE, int U, bool NT, bool BF>
__global__ void k_stream(const f4 *__restrict__ a, f4 *__restrict__ b, size_t n, int persistent,
                         float al, float be, float *sink) {
  const size_t per_block = (size_t)blockDim.x * U;
  const size_t nchunks = (n + per_block - 1) / per_block;
  f4 acc = {0, 0, 0, 0};
  for (size_t c = blockIdx.x; c < nchunks; c += (persistent ? gridDim.x : nchunks)) {
    const size_t base = c * per_block + threadIdx.x;
    f4 x[U], y[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = base + (size_t)u * blockDim.x;
      if (BF) {
        const size_t j = i < n ? i : 0;
        if (MODE == 0 || MODE == 2 || MODE == 3) x[u] = ld<NT>(a + j);
        if (MODE == 3) y[u] = ld<NT>(b + j);
      } else if (i < n) {
        if (MODE == 0 || MODE == 2 || MODE == 3) x[u] = ld<NT>(a + i);
        if (MODE == 3) y[u] = ld<NT>(b + i);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = base + (size_t)u * blockDim.x;
      if (i < n) {
        if (MODE == 0) acc += x[u];
        if (MODE == 1) st<NT>(b + i, f4{al, al, al, al});
        if (MODE == 2) st<NT>(b + i, x[u]);
        if (MODE == 3) st<NT>(b + i, be * y[u] + al * x[u]);
      }
    }
  }
  if (MODE == 0 && acc.x + acc.y + acc.z + acc.w == 1234.5f) sink[0] = acc.x;
}

typedef void (*KFn)(const f4 *, f4 *, size_t, int, float, float, float *);

template <int MODE, int U, bool NT>
KFn pick(bool bf) {
  return bf ? k_stream<MODE, U, NT, true> : k_stream<MODE, U, NT, false>;
}

static KFn table(int mode, int u, bool nt, bool bf) {
#define P(M)                                                                         \
  if (mode == M) {                                                                   \
    if (u == 1) return nt ? pick<M, 1, true>(bf) : pick<M, 1, false>(bf);            \
    if (u == 2) return nt ? pick<M, 2, true>(bf) : pick<M, 2, false>(bf);            \
    if (u == 4) return nt ? pick<M, 4, true>(bf) : pick<M, 4, false>(bf);            \
    if (u == 8) return nt ? pick<M, 8, true>(bf) : pick<M, 8, false>(bf);            \
  }
  P(0) P(1) P(2) P(3)
#undef P
  return nullptr;
}

int main(int argc, char **argv) {
  const size_t bytes = (argc > 1 ? std::atoll(argv[1]) : 2048) << 20;   // per buffer
  const size_t n = bytes / 16;
  f4 *a, *b;
  float *sink;
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  CK(hipMalloc(&sink, 64));
  CK(hipMemset(a, 0x3c, bytes));
  CK(hipMemset(b, 0x3c, bytes));
  int dev = 0, ncu = 0;
  CK(hipGetDevice(&dev));
  CK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const int reps = 20;
  auto timeit = [&](auto launch) {
    launch();
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r) launch();
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    return ms / reps;
  };
  // hipMemcpyAsync device to device (512 MiB and the full buffer)
  for (size_t sz : {(size_t)512 << 20, bytes}) {
    float ms = timeit([&] { CK(hipMemcpyAsync(b, a, sz, hipMemcpyDeviceToDevice, 0)); });
    std::printf("memcpyD2D bytes=%zu  %.1f GB/s (%.4f ms)\n", sz, 2.0 * sz / (ms * 1e-3) / 1e9, ms);
  }
  const char *names[] = {"read", "write", "copy", "axpby"};
  const double mult[] = {1, 1, 2, 3};
  for (int mode = 0; mode < 4; ++mode)
    for (int nt = 0; nt < 2; ++nt)
      for (int bf = 0; bf < 2; ++bf)
      for (int u : {1, 2, 4, 8})
        for (int blk : {256, 512, 1024})
          for (int pers : {0, 1}) {
            if (u * blk > 4096 && u > 2) continue;
            if (bf && u == 1) continue;   // one load per thread: no branch to remove
            KFn k = table(mode, u, nt, bf);
            const size_t per_block = (size_t)blk * u;
            const size_t nchunks = (n + per_block - 1) / per_block;
            const size_t grid = pers ? (size_t)ncu * (2048 / blk) * 2 : nchunks;
            float ms = timeit([&] {
              hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(blk), 0, 0, a, b, n, pers, 1.0f,
                                 0.5f, sink);
              CK(hipGetLastError());
            });
            std::printf("%-6s nt=%d bf=%d U=%d block=%4d %-10s %.1f GB/s\n", names[mode], nt, bf, u,
                        blk, pers ? "persistent" : "oneshot", mult[mode] * bytes / (ms * 1e-3) / 1e9);
            std::fflush(stdout);
          }
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
