// capi.cpp -- the extern "C" plan interface declared in include/hptt_amd.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "device.h"
#include "hptt_amd.h"
#include "kernels.h"
#include "planner.h"
#include "tune.h"

using namespace hpttx;

struct hpttx_plan {
  Problem prob;
  LaunchPlan shape;   // plan at the widest vectors the strides allow
  LaunchPlan cur;     // finalized for the current data pointers
  const void *A = nullptr;
  void *B = nullptr;
  hipStream_t stream = nullptr;
  CoefD coef{};
  bool update = false;
  bool cc = false;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

bool update_real(double beta) { return std::fabs(beta) >= 1e-16; }

// Run modes for the current pointers: a side is exact when the shape allows
// full 16-byte vectors and the pointer is 16-byte aligned, and unaligned
// (full-width vectors at element offsets, partial last vectors) otherwise;
// see finalize_vectors.
void refresh_vectors(hpttx_plan *p) {
  const uint32_t esz = p->prob.esize;
  const uint32_t cap = 16 / esz;
  uint32_t vr = p->shape.vr_shape, vw = p->shape.vw_shape;
  if (p->shape.kind == kRows) {
    uint32_t a = align_vec(p->A, esz, cap), b = align_vec(p->B, esz, cap);
    uint32_t v = vr;
    if (v > a || v > b) v = 1;
    if (v != cap) v = 1;
    p->cur = p->shape;
    finalize_vectors(p->prob, p->cur, v, v);
  } else {
    if (vr > align_vec(p->A, esz, cap) || vr != cap) vr = 1;
    if (vw > align_vec(p->B, esz, cap) || vw != cap) vw = 1;
    p->cur = p->shape;
    finalize_vectors(p->prob, p->cur, vr, vw);
    stream_for_pointers(p->prob, p->cur, p->A, p->B);
  }
}

int create_common(int dtype, const void *in, void *out, uint32_t rank,
                  const uint32_t *in_size, const uint32_t *perm, double ar,
                  double ai, double br, double bi, bool cc, double tuning_timeout,
                  const uint32_t *in_outer, const uint32_t *out_outer, void *stream,
                  hpttx_plan **plan) {
  if (!plan) return fail(HPTTX_EINVAL, "plan out-pointer is null");
  *plan = nullptr;
  // src/hptt.cc:27-28: null data pointers -> nullptr.
  if (!in || !out) return fail(HPTTX_EINVAL, "null data pointer");
  hpttx_plan *p = new (std::nothrow) hpttx_plan();
  if (!p) return fail(HPTTX_ENOMEM, "out of host memory");
  std::string err;
  int rc = build_problem(dtype, rank, in_size, perm, in_outer, out_outer, p->prob, err);
  if (rc != kOk) {
    delete p;
    return fail(rc, err);
  }
  // Arguments first, then the device: the reference validates before it
  // looks for a plugin that fits the host and returns nullptr when none does
  // (src/hptt.cc:25-56, 86-87; src/arch/arch.cc:74-89).
  rc = require_target_device(err);
  if (rc != kOk) {
    delete p;
    return fail(rc, err);
  }
  p->A = in;
  p->B = out;
  p->stream = static_cast<hipStream_t>(stream);
  p->cc = cc && (ai != 0.0 || bi != 0.0);
  p->coef = CoefD{ar, p->cc ? ai : 0.0, br, p->cc ? bi : 0.0};
  p->update = p->cc ? (std::fabs(br) >= 1e-16 || std::fabs(bi) >= 1e-16) : update_real(br);
  TuneParams tp{};
  const bool cached = tune_lookup(p->prob, p->update, p->cc, tp);
  if (!cached && tuning_timeout != 0.0) {
    rc = autotune(p->prob, p->coef, p->update, p->cc, tuning_timeout, p->stream, tp, err, in, out);
    if (rc != kOk) {
      delete p;
      return fail(rc, err);
    }
  }
  rc = plan_launch(p->prob, tp, p->shape, err, kMinPlanTiles);
  if (rc != kOk && cached) {
    // A cached record that does not plan (from another build or a damaged
    // file) is replaced by the model's pick rather than failing the plan.
    tune_forget(tune_key(p->prob, p->update, p->cc));
    rc = plan_launch(p->prob, TuneParams{}, p->shape, err, kMinPlanTiles);
  }
  if (rc != kOk) {
    delete p;
    return fail(rc, err);
  }
  refresh_vectors(p);
  *plan = p;
  g_err.clear();
  return HPTTX_OK;
}

}  // namespace

namespace hpttx {
void set_last_error(const std::string &m) { g_err = m; }
}  // namespace hpttx

extern "C" {

int hpttx_plan_create_s(const float *in, float *out, uint32_t rank, const uint32_t *in_size,
                        const uint32_t *perm, float alpha, float beta, uint32_t,
                        double tuning_timeout, const uint32_t *in_outer,
                        const uint32_t *out_outer, void *stream, hpttx_plan **plan) {
  return create_common(kF32, in, out, rank, in_size, perm, alpha, 0.0, beta, 0.0, false,
                       tuning_timeout, in_outer, out_outer, stream, plan);
}

int hpttx_plan_create_d(const double *in, double *out, uint32_t rank, const uint32_t *in_size,
                        const uint32_t *perm, double alpha, double beta, uint32_t,
                        double tuning_timeout, const uint32_t *in_outer,
                        const uint32_t *out_outer, void *stream, hpttx_plan **plan) {
  return create_common(kF64, in, out, rank, in_size, perm, alpha, 0.0, beta, 0.0, false,
                       tuning_timeout, in_outer, out_outer, stream, plan);
}

int hpttx_plan_create_c(const void *in, void *out, uint32_t rank, const uint32_t *in_size,
                        const uint32_t *perm, float alpha, float beta, uint32_t,
                        double tuning_timeout, const uint32_t *in_outer,
                        const uint32_t *out_outer, void *stream, hpttx_plan **plan) {
  return create_common(kC64, in, out, rank, in_size, perm, alpha, 0.0, beta, 0.0, false,
                       tuning_timeout, in_outer, out_outer, stream, plan);
}

int hpttx_plan_create_z(const void *in, void *out, uint32_t rank, const uint32_t *in_size,
                        const uint32_t *perm, double alpha, double beta, uint32_t,
                        double tuning_timeout, const uint32_t *in_outer,
                        const uint32_t *out_outer, void *stream, hpttx_plan **plan) {
  return create_common(kC128, in, out, rank, in_size, perm, alpha, 0.0, beta, 0.0, false,
                       tuning_timeout, in_outer, out_outer, stream, plan);
}

int hpttx_plan_create_c_cplx(const void *in, void *out, uint32_t rank,
                             const uint32_t *in_size, const uint32_t *perm, float alpha_re,
                             float alpha_im, float beta_re, float beta_im, uint32_t,
                             double tuning_timeout, const uint32_t *in_outer,
                             const uint32_t *out_outer, void *stream, hpttx_plan **plan) {
  return create_common(kC64, in, out, rank, in_size, perm, alpha_re, alpha_im, beta_re,
                       beta_im, true, tuning_timeout, in_outer, out_outer, stream, plan);
}

int hpttx_plan_create_z_cplx(const void *in, void *out, uint32_t rank,
                             const uint32_t *in_size, const uint32_t *perm, double alpha_re,
                             double alpha_im, double beta_re, double beta_im, uint32_t,
                             double tuning_timeout, const uint32_t *in_outer,
                             const uint32_t *out_outer, void *stream, hpttx_plan **plan) {
  return create_common(kC128, in, out, rank, in_size, perm, alpha_re, alpha_im, beta_re,
                       beta_im, true, tuning_timeout, in_outer, out_outer, stream, plan);
}

int hpttx_plan_exec_on(hpttx_plan *p, void *stream) {
  if (!p) return fail(HPTTX_EINVAL, "null plan");
  hipError_t e = launch_plan(p->prob.dtype, p->cur, p->A, p->B, p->coef, p->update, p->cc,
                             static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(HPTTX_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return HPTTX_OK;
}

int hpttx_plan_exec(hpttx_plan *p) {
  if (!p) return fail(HPTTX_EINVAL, "null plan");
  return hpttx_plan_exec_on(p, p->stream);
}

int hpttx_plan_sync(hpttx_plan *p) {
  if (!p) return fail(HPTTX_EINVAL, "null plan");
  hipError_t e = hipStreamSynchronize(p->stream);
  if (e != hipSuccess) return fail(HPTTX_EHIP, std::string("stream sync: ") + hipGetErrorString(e));
  return HPTTX_OK;
}

int hpttx_plan_reset_data(hpttx_plan *p, const void *in, void *out) {
  if (!p) return fail(HPTTX_EINVAL, "null plan");
  if (!in || !out) return fail(HPTTX_EINVAL, "null data pointer");
  p->A = in;
  p->B = out;
  refresh_vectors(p);
  return HPTTX_OK;
}

int hpttx_plan_set_stream(hpttx_plan *p, void *stream) {
  if (!p) return fail(HPTTX_EINVAL, "null plan");
  p->stream = static_cast<hipStream_t>(stream);
  return HPTTX_OK;
}

int hpttx_plan_describe(const hpttx_plan *p, char *buf, size_t len) {
  if (!p) return fail(HPTTX_EINVAL, "null plan");
  std::string s = describe(p->prob, p->cur);
  s += p->update ? "update: B is read (beta*B + alpha*A)\n" : "update: no (B = alpha*A, B not read)\n";
  if (buf && len) {
    size_t n = s.size() < len - 1 ? s.size() : len - 1;
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
  }
  return (int)s.size() + 1;
}

int hpttx_plan_get_info(const hpttx_plan *p, hpttx_plan_info *info) {
  if (!p || !info) return fail(HPTTX_EINVAL, "null argument");
  std::memset(info, 0, sizeof(*info));
  info->elements = p->prob.N;
  info->elem_bytes = p->prob.esize;
  info->update = p->update ? 1 : 0;
  const std::string &path = p->cur.path;
  info->path = path == "tile" ? 0 : path == "tile-shared" ? 1 : path == "wide-copy" ? 2
             : path == "axpby" ? 3 : 4;
  info->fused_rank = p->prob.fr;
  info->algorithmic_bytes = p->prob.N * p->prob.esize * (uint64_t)(2 + (p->update ? 1 : 0));
  info->grid = p->cur.grid;
  info->lds_bytes = (uint32_t)p->cur.lds_bytes;
  info->threads = p->cur.threads;
  if (p->cur.kind == kTile) {
    info->tile_r = p->cur.tile.TR * p->cur.tile.VS;
    info->tile_c = p->cur.tile.TC * p->cur.tile.VS;
    info->tile_o = p->cur.tile.TO;
    info->vec_r = p->cur.tile.vr;
    info->vec_w = p->cur.tile.vw;
    info->unaligned_r = p->cur.tile.ua_r;
    info->unaligned_w = p->cur.tile.ua_w;
  } else {
    info->tile_r = p->cur.rows.SEG;
    info->tile_c = 1;
    info->tile_o = p->cur.rows.RB;
    info->vec_r = info->vec_w = p->cur.rows.v;
    info->unaligned_r = info->unaligned_w = p->cur.rows.ua;
  }
  return HPTTX_OK;
}

void hpttx_plan_destroy(hpttx_plan *p) { delete p; }

const char *hpttx_last_error(void) { return g_err.c_str(); }

int hpttx_version(void) { return HPTTX_VERSION; }

}  // extern "C"
