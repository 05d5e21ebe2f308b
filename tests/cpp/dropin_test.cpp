// TEST INFRASTRUCTURE: the drop-in C++ interface (include/hptt/hptt.h) used
// the way the reference's README uses hptt::create_plan (README.md:53-70),
// on device buffers, checked bit for bit against the oracle (liboracle.so).
#include <hip/hip_runtime.h>

#include <complex>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "hptt/hptt.h"

extern "C" int oracle_transpose(int dtype, const void *A, void *B, uint32_t rank,
                                const uint32_t *in_size, const uint32_t *perm, double ar,
                                double ai, double br, double bi, int cplx_coef,
                                const uint32_t *in_outer, const uint32_t *out_outer,
                                int nthreads);

static int failures = 0;
#define EXPECT(c)                                                  \
  do {                                                             \
    if (!(c)) {                                                    \
      std::printf("FAIL line %d: %s\n", __LINE__, #c);             \
      ++failures;                                                  \
    }                                                              \
  } while (0)

template <typename T>
static void fill(std::vector<T> &v, unsigned seed) {
  std::mt19937 g(seed);
  std::uniform_real_distribution<double> d(-1.0, 1.0);
  for (auto &x : v) {
    if constexpr (sizeof(T) == 4) x = (T)d(g);
    else if constexpr (std::is_same<T, double>::value) x = d(g);
    else {
      using R = decltype(std::real(T()));
      x = T((R)d(g), (R)d(g));
    }
  }
}

template <typename T>
static T *to_device(const std::vector<T> &h) {
  T *d = nullptr;
  if (hipMalloc(&d, h.size() * sizeof(T)) != hipSuccess) return nullptr;
  (void)hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
  return d;
}

template <typename T>
static std::vector<T> to_host(const T *d, size_t n) {
  std::vector<T> h(n);
  (void)hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost);
  return h;
}

template <typename T, typename Coef>
static void case_real(int code, std::vector<uint32_t> size, std::vector<uint32_t> perm, Coef a,
                      Coef b) {
  size_t n = 1;
  for (auto s : size) n *= s;
  std::vector<T> A(n), B(n);
  fill(A, 1);
  fill(B, 2);
  T *dA = to_device(A), *dB = to_device(B);
  // README.md:60-63: create_plan(in, out, in_size, perm, 2.3f, 4.2f, 20); exec
  auto plan = hptt::create_plan(dA, dB, size, perm, a, b, 20);
  EXPECT(plan != nullptr);
  if (!plan) return;
  plan->exec();
  plan->sync();
  std::vector<T> got = to_host(dB, n);
  std::vector<T> exp = B;
  EXPECT(oracle_transpose(code, A.data(), exp.data(), (uint32_t)perm.size(), size.data(),
                          perm.data(), (double)a, 0.0, (double)b, 0.0, 0, nullptr, nullptr,
                          4) == 0);
  EXPECT(std::memcmp(got.data(), exp.data(), n * sizeof(T)) == 0);
  // README.md:66-69: reuse the plan on new buffers of the same shape.
  std::vector<T> A2(n), B2(n);
  fill(A2, 3);
  fill(B2, 4);
  T *dA2 = to_device(A2), *dB2 = to_device(B2);
  plan->reset_data(dA2, dB2);
  (*plan)();
  plan->sync();
  got = to_host(dB2, n);
  exp = B2;
  oracle_transpose(code, A2.data(), exp.data(), (uint32_t)perm.size(), size.data(), perm.data(),
                   (double)a, 0.0, (double)b, 0.0, 0, nullptr, nullptr, 4);
  EXPECT(std::memcmp(got.data(), exp.data(), n * sizeof(T)) == 0);
  plan->set_thread_ids({0, 1});
  plan->unset_thread_ids();
  EXPECT(plan->describe().find("path:") != std::string::npos);
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dA2);
  (void)hipFree(dB2);
}

int main() {
  case_real<float>(0, {40, 30, 20}, {2, 1, 0}, 2.3f, 4.2f);
  case_real<double>(1, {16, 12, 8, 4, 6, 8}, {5, 2, 0, 4, 1, 3}, 2.3, 4.2);
  case_real<double>(1, {33, 17, 5}, {0, 2, 1}, 1.0, 0.0);

  // complex<double> with complex coefficients on a stream
  {
    std::vector<uint32_t> size{24, 20, 18, 16}, perm{3, 2, 1, 0};
    size_t n = 24 * 20 * 18 * 16;
    std::vector<std::complex<double>> A(n), B(n);
    fill(A, 5);
    fill(B, 6);
    auto *dA = to_device(A), *dB = to_device(B);
    hipStream_t st;
    (void)hipStreamCreate(&st);
    auto plan = hptt::create_plan(dA, dB, size, perm, std::complex<double>(1.1, 0.3),
                                  std::complex<double>(0.7, -0.2), 0, 0.0, {}, {}, st);
    EXPECT(plan != nullptr);
    if (plan) {
      plan->exec();
      plan->sync();
      auto got = to_host(dB, n);
      auto exp = B;
      oracle_transpose(3, A.data(), exp.data(), 4, size.data(), perm.data(), 1.1, 0.3, 0.7, -0.2,
                       1, nullptr, nullptr, 4);
      EXPECT(std::memcmp(got.data(), exp.data(), n * sizeof(exp[0])) == 0);
    }
    (void)hipStreamDestroy(st);
    (void)hipFree(dA);
    (void)hipFree(dB);
  }

  // float _Complex with real coefficients (the reference's own types)
  {
    std::vector<uint32_t> size{10, 9, 8}, perm{1, 2, 0};
    size_t n = 720;
    std::vector<std::complex<float>> A(n), B(n);
    fill(A, 7);
    fill(B, 8);
    auto *dA = to_device(A), *dB = to_device(B);
    auto *cA = reinterpret_cast<hptt::FloatComplex *>(dA);
    auto *cB = reinterpret_cast<hptt::FloatComplex *>(dB);
    auto plan = hptt::create_plan(cA, cB, size, perm, 2.0f, -1.5f, 4);
    EXPECT(plan != nullptr);
    if (plan) {
      plan->exec();
      plan->sync();
      auto got = to_host(dB, n);
      auto exp = B;
      oracle_transpose(2, A.data(), exp.data(), 3, size.data(), perm.data(), 2.0, 0.0, -1.5, 0.0,
                       0, nullptr, nullptr, 4);
      EXPECT(std::memcmp(got.data(), exp.data(), n * sizeof(exp[0])) == 0);
    }
    (void)hipFree(dA);
    (void)hipFree(dB);
  }

  // nullptr cases (inc/hptt/hptt.h:71-83)
  {
    float *d = nullptr;
    (void)hipMalloc(&d, 1024);
    EXPECT(hptt::create_plan<float>(nullptr, d, {4, 4}, {1, 0}, 1.f, 0.f, 1) == nullptr);
    EXPECT(hptt::create_plan<float>(d, d + 16, {4}, {0}, 1.f, 0.f, 1) == nullptr);
    EXPECT(hptt::create_plan<float>(d, d + 16, {4, 4}, {1, 1}, 1.f, 0.f, 1) == nullptr);
    EXPECT(hptt::create_plan<float>(d, d + 16, {4, 4, 2}, {1, 0}, 1.f, 0.f, 1) == nullptr);
    EXPECT(hptt::create_plan<float>(d, d + 16, {4, 0}, {1, 0}, 1.f, 0.f, 1) == nullptr);
    EXPECT(hptt::create_plan<float>(d, d + 16, {4, 4}, {1, 0}, 1.f, 0.f, 1, 0.0, {3, 4}) == nullptr);
    EXPECT(hptt::create_plan<float>(d, d + 16, {2, 2, 2, 2, 2, 2, 2}, {0, 1, 2, 3, 4, 5, 6}, 1.f,
                                    0.f, 1) == nullptr);
    EXPECT(hptt::create_plan<float>(d, d + 16, {4, 4}, {1, 0}, 1.f, 0.f, 1) != nullptr);
    (void)hipFree(d);
  }
  if (failures) {
    std::printf("DROPIN FAILED (%d)\n", failures);
    return 1;
  }
  std::printf("DROPIN OK\n");
  return 0;
}
