// hptt/hptt.h -- drop-in C++ interface of the MI355X transposition library.
//
// Same names, signatures and failure behaviour as the reference's public
// header (inc/hptt/hptt.h:85-123, inc/hptt/impl/hptt_trans.h:12-36,
// inc/hptt/types.h:14-54):
//
//   auto plan = hptt::create_plan(in, out, in_size, perm, alpha, beta, threads);
//   if (plan) plan->exec();
//   plan->reset_data(in2, out2);
//
// Differences a caller must know:
//   * in / out are DEVICE pointers (hipMalloc'd or any device-accessible
//     allocation), tensors column-major as in the reference;
//   * exec() / operator() enqueue on the plan's HIP stream and return; call
//     sync() (or synchronise the stream) before reading B on the host;
//   * num_threads is accepted and ignored; set_thread_ids / unset_thread_ids
//     are accepted no-ops (there are no host threads to pin);
//   * rank > 6 returns nullptr (the reference returns a plan that does
//     nothing); the out_outer_size check compares B's outer size with B's own
//     extent (in_size[perm[d]]).
// Added: overloads taking the stream, complex-coefficient overloads for the
// complex types, sync(), describe(), native_handle(), status().  exec() does
// not throw: a failed launch is written to stderr and reported by status()
// and hpttx_last_error().
//
// Header-only over the C ABI in hptt_amd.h; link with -lhptt_amd.
#pragma once
#ifndef HPTT_AMD_HPTT_H_
#define HPTT_AMD_HPTT_H_

#include <complex>
#include <cstdio>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "hptt_amd.h"

namespace hptt {

// inc/hptt/types.h:14-54
using TensorIdx = uint64_t;
using TensorInt = int32_t;
using TensorUInt = uint32_t;
using FloatComplex = float _Complex;
using DoubleComplex = double _Complex;

template <typename FloatType> struct FloatTypeDeducer {};
template <> struct FloatTypeDeducer<float> { using type = float; };
template <> struct FloatTypeDeducer<double> { using type = double; };
template <> struct FloatTypeDeducer<FloatComplex> { using type = float; };
template <> struct FloatTypeDeducer<DoubleComplex> { using type = double; };
template <> struct FloatTypeDeducer<std::complex<float>> { using type = float; };
template <> struct FloatTypeDeducer<std::complex<double>> { using type = double; };
template <typename FloatType>
using DeducedFloatType = typename FloatTypeDeducer<FloatType>::type;

// inc/hptt/impl/hptt_trans.h:12-36, with exec() asynchronous on a stream.
template <typename FloatType>
class CGraphTransPackBase {
 public:
  CGraphTransPackBase() = default;
  CGraphTransPackBase(const CGraphTransPackBase &) = delete;
  CGraphTransPackBase &operator=(const CGraphTransPackBase &) = delete;
  CGraphTransPackBase(CGraphTransPackBase &&) = delete;
  CGraphTransPackBase &operator=(CGraphTransPackBase &&) = delete;
  virtual ~CGraphTransPackBase() = default;

  virtual void exec() = 0;
  virtual void operator()() = 0;
  virtual void print_plan() = 0;
  virtual void reset_data(const FloatType *in_data, FloatType *out_data) = 0;
  virtual void set_thread_ids(const std::vector<int32_t> &thread_ids) = 0;
  virtual void unset_thread_ids() = 0;
  // Added: wait for the plan's stream; human-readable plan; C handle; the
  // status (HPTTX_*) of the last exec / reset_data / sync.
  virtual void sync() = 0;
  virtual std::string describe() const = 0;
  virtual hpttx_plan *native_handle() const = 0;
  virtual int status() const = 0;
};

template <typename FloatType>
using trans_plan = CGraphTransPackBase<FloatType>;

namespace detail {

template <typename FloatType>
class DevicePlan final : public CGraphTransPackBase<FloatType> {
 public:
  explicit DevicePlan(hpttx_plan *p) : p_(p) {}
  ~DevicePlan() override { hpttx_plan_destroy(p_); }
  // The reference's exec() / reset_data() return void and never fail
  // (inc/hptt/impl/hptt_trans.h:27-30); a GPU launch can.  They stay void
  // and do not throw: a failure is written to stderr, kept as status() and
  // in hpttx_last_error().
  void exec() override { report(hpttx_plan_exec(p_), "exec"); }
  void operator()() override { exec(); }
  void print_plan() override {
    const std::string s = describe();
    std::fputs(s.c_str(), stdout);
  }
  void reset_data(const FloatType *in_data, FloatType *out_data) override {
    report(hpttx_plan_reset_data(p_, in_data, out_data), "reset_data");
  }
  void set_thread_ids(const std::vector<int32_t> &) override {}
  void unset_thread_ids() override {}
  void sync() override { report(hpttx_plan_sync(p_), "sync"); }
  int status() const override { return status_; }
  std::string describe() const override {
    const int n = hpttx_plan_describe(p_, nullptr, 0);
    std::string s(n > 0 ? (size_t)n : 1, '\0');
    hpttx_plan_describe(p_, &s[0], s.size());
    s.resize(s.find('\0') == std::string::npos ? s.size() : s.find('\0'));
    return s;
  }
  hpttx_plan *native_handle() const override { return p_; }

 private:
  void report(int rc, const char *what) {
    status_ = rc;
    if (rc != HPTTX_OK) std::fprintf(stderr, "hptt_amd: %s failed: %s\n", what, hpttx_last_error());
  }
  hpttx_plan *p_;
  int status_ = HPTTX_OK;
};

inline const uint32_t *opt(const std::vector<uint32_t> &v) { return v.empty() ? nullptr : v.data(); }

// Sizes must agree before the C ABI sees the arrays (src/hptt.cc:35-39).
inline bool sizes_ok(const std::vector<uint32_t> &in_size, const std::vector<uint32_t> &perm,
                     const std::vector<uint32_t> &io, const std::vector<uint32_t> &oo) {
  return perm.size() == in_size.size() && (io.empty() || io.size() == perm.size()) &&
         (oo.empty() || oo.size() == perm.size());
}

template <typename FloatType>
std::shared_ptr<trans_plan<FloatType>> wrap(int rc, hpttx_plan *p) {
  if (rc != HPTTX_OK || !p) return nullptr;
  return std::shared_ptr<trans_plan<FloatType>>(new DevicePlan<FloatType>(p));
}

inline int create(const float *in, float *out, uint32_t r, const uint32_t *s, const uint32_t *pm,
                  float a, float b, uint32_t t, double to, const uint32_t *io, const uint32_t *oo,
                  void *st, hpttx_plan **p) {
  return hpttx_plan_create_s(in, out, r, s, pm, a, b, t, to, io, oo, st, p);
}
inline int create(const double *in, double *out, uint32_t r, const uint32_t *s, const uint32_t *pm,
                  double a, double b, uint32_t t, double to, const uint32_t *io, const uint32_t *oo,
                  void *st, hpttx_plan **p) {
  return hpttx_plan_create_d(in, out, r, s, pm, a, b, t, to, io, oo, st, p);
}
inline int create(const FloatComplex *in, FloatComplex *out, uint32_t r, const uint32_t *s,
                  const uint32_t *pm, float a, float b, uint32_t t, double to, const uint32_t *io,
                  const uint32_t *oo, void *st, hpttx_plan **p) {
  return hpttx_plan_create_c(in, out, r, s, pm, a, b, t, to, io, oo, st, p);
}
inline int create(const DoubleComplex *in, DoubleComplex *out, uint32_t r, const uint32_t *s,
                  const uint32_t *pm, double a, double b, uint32_t t, double to,
                  const uint32_t *io, const uint32_t *oo, void *st, hpttx_plan **p) {
  return hpttx_plan_create_z(in, out, r, s, pm, a, b, t, to, io, oo, st, p);
}
inline int create(const std::complex<float> *in, std::complex<float> *out, uint32_t r,
                  const uint32_t *s, const uint32_t *pm, float a, float b, uint32_t t, double to,
                  const uint32_t *io, const uint32_t *oo, void *st, hpttx_plan **p) {
  return hpttx_plan_create_c(in, out, r, s, pm, a, b, t, to, io, oo, st, p);
}
inline int create(const std::complex<double> *in, std::complex<double> *out, uint32_t r,
                  const uint32_t *s, const uint32_t *pm, double a, double b, uint32_t t,
                  double to, const uint32_t *io, const uint32_t *oo, void *st, hpttx_plan **p) {
  return hpttx_plan_create_z(in, out, r, s, pm, a, b, t, to, io, oo, st, p);
}

}  // namespace detail

// inc/hptt/hptt.h:85-93.  nullptr on invalid arguments (inc/hptt/hptt.h:71-83).
template <typename FloatType>
std::shared_ptr<trans_plan<FloatType>> create_plan(
    const FloatType *in_data, FloatType *out_data, const std::vector<uint32_t> &in_size,
    const std::vector<uint32_t> &perm, const DeducedFloatType<FloatType> alpha,
    const DeducedFloatType<FloatType> beta, const uint32_t num_threads,
    const double tuning_timeout = 0.0, const std::vector<uint32_t> &in_outer_size = {},
    const std::vector<uint32_t> &out_outer_size = {}) {
  if (!detail::sizes_ok(in_size, perm, in_outer_size, out_outer_size)) return nullptr;
  hpttx_plan *p = nullptr;
  const int rc = detail::create(in_data, out_data, (uint32_t)perm.size(), in_size.data(),
                                perm.data(), alpha, beta, num_threads, tuning_timeout,
                                detail::opt(in_outer_size), detail::opt(out_outer_size), nullptr,
                                &p);
  return detail::wrap<FloatType>(rc, p);
}

// Added: the same, enqueueing on `stream` (a hipStream_t).
template <typename FloatType>
std::shared_ptr<trans_plan<FloatType>> create_plan(
    const FloatType *in_data, FloatType *out_data, const std::vector<uint32_t> &in_size,
    const std::vector<uint32_t> &perm, const DeducedFloatType<FloatType> alpha,
    const DeducedFloatType<FloatType> beta, const uint32_t num_threads,
    const double tuning_timeout, const std::vector<uint32_t> &in_outer_size,
    const std::vector<uint32_t> &out_outer_size, void *stream) {
  if (!detail::sizes_ok(in_size, perm, in_outer_size, out_outer_size)) return nullptr;
  hpttx_plan *p = nullptr;
  const int rc = detail::create(in_data, out_data, (uint32_t)perm.size(), in_size.data(),
                                perm.data(), alpha, beta, num_threads, tuning_timeout,
                                detail::opt(in_outer_size), detail::opt(out_outer_size), stream,
                                &p);
  return detail::wrap<FloatType>(rc, p);
}

// Added: complex coefficients for the complex element types (fixed order,
// see hpttx_plan_create_c_cplx).
inline std::shared_ptr<trans_plan<std::complex<float>>> create_plan(
    const std::complex<float> *in_data, std::complex<float> *out_data,
    const std::vector<uint32_t> &in_size, const std::vector<uint32_t> &perm,
    std::complex<float> alpha, std::complex<float> beta, const uint32_t num_threads,
    const double tuning_timeout = 0.0, const std::vector<uint32_t> &in_outer_size = {},
    const std::vector<uint32_t> &out_outer_size = {}, void *stream = nullptr) {
  if (!detail::sizes_ok(in_size, perm, in_outer_size, out_outer_size)) return nullptr;
  hpttx_plan *p = nullptr;
  const int rc = hpttx_plan_create_c_cplx(
      in_data, out_data, (uint32_t)perm.size(), in_size.data(), perm.data(), alpha.real(),
      alpha.imag(), beta.real(), beta.imag(), num_threads, tuning_timeout,
      detail::opt(in_outer_size), detail::opt(out_outer_size), stream, &p);
  return detail::wrap<std::complex<float>>(rc, p);
}

inline std::shared_ptr<trans_plan<std::complex<double>>> create_plan(
    const std::complex<double> *in_data, std::complex<double> *out_data,
    const std::vector<uint32_t> &in_size, const std::vector<uint32_t> &perm,
    std::complex<double> alpha, std::complex<double> beta, const uint32_t num_threads,
    const double tuning_timeout = 0.0, const std::vector<uint32_t> &in_outer_size = {},
    const std::vector<uint32_t> &out_outer_size = {}, void *stream = nullptr) {
  if (!detail::sizes_ok(in_size, perm, in_outer_size, out_outer_size)) return nullptr;
  hpttx_plan *p = nullptr;
  const int rc = hpttx_plan_create_z_cplx(
      in_data, out_data, (uint32_t)perm.size(), in_size.data(), perm.data(), alpha.real(),
      alpha.imag(), beta.real(), beta.imag(), num_threads, tuning_timeout,
      detail::opt(in_outer_size), detail::opt(out_outer_size), stream, &p);
  return detail::wrap<std::complex<double>>(rc, p);
}

}  // namespace hptt

#endif  // HPTT_AMD_HPTT_H_
