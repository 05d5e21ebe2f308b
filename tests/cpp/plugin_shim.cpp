// plugin_shim.cpp -- the reference-side binding for a maintainer who keeps
// the reference's own front end (src/hptt.cc) and swaps its per-ISA plugin.
//
// The reference's hptt::create_plan<T> dlopens a plugin library and looks up
// four C++ factories by their mangled names (src/hptt.cc:66-92), each
// returning a raw CGraphTransPackBase<T>* that it wraps in a shared_ptr
// (signatures: src/impl/hptt_trans_impl_s.cc:8-16 and the d/c/z twins).
// This file defines those four factories over the C ABI of include/hptt_amd.h,
// so a library built from it, installed under the plugin's name, serves the
// unchanged reference front end with MI355X plans.  exec() keeps the
// reference's synchronous meaning (enqueue, then wait); the data pointers
// must be device pointers.
//
// Compiled against the reference's public headers (inc/hptt/types.h,
// inc/hptt/impl/hptt_trans.h); tests/test_capi.py builds it and checks the
// exported names against the ones src/hptt.cc looks up.
#include <cstdio>
#include <vector>

#include <hptt/impl/hptt_trans.h>
#include <hptt/types.h>

#include "hptt_amd.h"

namespace {

template <typename T>
class HpttxPlan final : public hptt::CGraphTransPackBase<T> {
 public:
  explicit HpttxPlan(hpttx_plan *p) : p_(p) {}
  ~HpttxPlan() override { hpttx_plan_destroy(p_); }
  // The reference's exec() returns void and cannot report failure
  // (inc/hptt/impl/hptt_trans.h:27): a failed launch or wait is written to
  // stderr and stays readable through hpttx_last_error().
  void exec() override {
    int rc = hpttx_plan_exec(p_);
    if (rc == HPTTX_OK) rc = hpttx_plan_sync(p_);
    if (rc != HPTTX_OK) std::fprintf(stderr, "hptt plugin: exec failed: %s\n", hpttx_last_error());
  }
  void operator()() override { exec(); }
  void print_plan() override {
    char buf[4096];
    (void)hpttx_plan_describe(p_, buf, sizeof buf);
    std::fputs(buf, stdout);
  }
  void reset_data(const T *in_data, T *out_data) override {
    if (hpttx_plan_reset_data(p_, in_data, out_data) != HPTTX_OK)
      std::fprintf(stderr, "hptt plugin: reset_data failed: %s\n", hpttx_last_error());
  }
  void set_thread_ids(const std::vector<int32_t> &) override {}
  void unset_thread_ids() override {}

 private:
  hpttx_plan *p_;
};

const uint32_t *opt(const std::vector<hptt::TensorUInt> &v) {
  return v.empty() ? nullptr : v.data();
}

// The reference checks sizes before calling a factory (src/hptt.cc:35-39);
// the C ABI checks the values.
template <typename T>
hptt::CGraphTransPackBase<T> *wrap(int rc, hpttx_plan *p) {
  if (rc != HPTTX_OK || !p) return nullptr;
  return new HpttxPlan<T>(p);
}

}  // namespace

hptt::CGraphTransPackBase<float> *create_trans_plan_impl_s(
    const float *in_data, float *out_data, const hptt::TensorUInt order,
    const std::vector<hptt::TensorUInt> &in_size, const std::vector<hptt::TensorUInt> &perm,
    const hptt::DeducedFloatType<float> alpha, const hptt::DeducedFloatType<float> beta,
    const hptt::TensorUInt num_threads, const double tuning_timeout,
    const std::vector<hptt::TensorUInt> &in_outer_size,
    const std::vector<hptt::TensorUInt> &out_outer_size) {
  hpttx_plan *p = nullptr;
  const int rc = hpttx_plan_create_s(in_data, out_data, order, in_size.data(), perm.data(), alpha,
                                     beta, num_threads, tuning_timeout, opt(in_outer_size),
                                     opt(out_outer_size), nullptr, &p);
  return wrap<float>(rc, p);
}

hptt::CGraphTransPackBase<double> *create_trans_plan_impl_d(
    const double *in_data, double *out_data, const hptt::TensorUInt order,
    const std::vector<hptt::TensorUInt> &in_size, const std::vector<hptt::TensorUInt> &perm,
    const hptt::DeducedFloatType<double> alpha, const hptt::DeducedFloatType<double> beta,
    const hptt::TensorUInt num_threads, const double tuning_timeout,
    const std::vector<hptt::TensorUInt> &in_outer_size,
    const std::vector<hptt::TensorUInt> &out_outer_size) {
  hpttx_plan *p = nullptr;
  const int rc = hpttx_plan_create_d(in_data, out_data, order, in_size.data(), perm.data(), alpha,
                                     beta, num_threads, tuning_timeout, opt(in_outer_size),
                                     opt(out_outer_size), nullptr, &p);
  return wrap<double>(rc, p);
}

hptt::CGraphTransPackBase<hptt::FloatComplex> *create_trans_plan_impl_c(
    const hptt::FloatComplex *in_data, hptt::FloatComplex *out_data,
    const hptt::TensorUInt order, const std::vector<hptt::TensorUInt> &in_size,
    const std::vector<hptt::TensorUInt> &perm,
    const hptt::DeducedFloatType<hptt::FloatComplex> alpha,
    const hptt::DeducedFloatType<hptt::FloatComplex> beta, const hptt::TensorUInt num_threads,
    const double tuning_timeout, const std::vector<hptt::TensorUInt> &in_outer_size,
    const std::vector<hptt::TensorUInt> &out_outer_size) {
  hpttx_plan *p = nullptr;
  const int rc = hpttx_plan_create_c(in_data, out_data, order, in_size.data(), perm.data(), alpha,
                                     beta, num_threads, tuning_timeout, opt(in_outer_size),
                                     opt(out_outer_size), nullptr, &p);
  return wrap<hptt::FloatComplex>(rc, p);
}

hptt::CGraphTransPackBase<hptt::DoubleComplex> *create_trans_plan_impl_z(
    const hptt::DoubleComplex *in_data, hptt::DoubleComplex *out_data,
    const hptt::TensorUInt order, const std::vector<hptt::TensorUInt> &in_size,
    const std::vector<hptt::TensorUInt> &perm,
    const hptt::DeducedFloatType<hptt::DoubleComplex> alpha,
    const hptt::DeducedFloatType<hptt::DoubleComplex> beta, const hptt::TensorUInt num_threads,
    const double tuning_timeout, const std::vector<hptt::TensorUInt> &in_outer_size,
    const std::vector<hptt::TensorUInt> &out_outer_size) {
  hpttx_plan *p = nullptr;
  const int rc = hpttx_plan_create_z(in_data, out_data, order, in_size.data(), perm.data(), alpha,
                                     beta, num_threads, tuning_timeout, opt(in_outer_size),
                                     opt(out_outer_size), nullptr, &p);
  return wrap<hptt::DoubleComplex>(rc, p);
}
