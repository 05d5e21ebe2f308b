// batch_capi.cpp -- the extern "C" batch interface declared in include/hptt_amd.h.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "batch.h"
#include "hptt_amd.h"

using namespace hpttx;

struct hpttx_batch {
  Batch b;
};

// Shares hpttx_last_error with the plan interface (capi.cpp).
namespace hpttx {
void set_last_error(const std::string &m);

// Validate one C item and convert it.  Coefficients are held at the
// element's precision (float for F32 / C64), as the plan interface does.
int batch_item_to_spec(const hpttx_batch_item &it, uint32_t k, BatchItemSpec &s,
                       bool allow_null_out) {
  std::memset(&s, 0, sizeof(s));
  const std::string where = "item " + std::to_string(k) + ": ";
  if (it.dtype < HPTTX_F32 || it.dtype > HPTTX_C128) {
    set_last_error(where + "unknown dtype");
    return HPTTX_EINVAL;
  }
  if (it.rank > 6) {
    set_last_error(where + "rank above 6 is not supported");
    return HPTTX_EUNSUPPORTED;
  }
  if (!it.in || (!it.out && !allow_null_out)) {
    set_last_error(where + "null data pointer");
    return HPTTX_EINVAL;
  }
  const bool cplx_type = it.dtype == HPTTX_C64 || it.dtype == HPTTX_C128;
  if (it.complex_coef && !cplx_type) {
    set_last_error(where + "complex coefficients need a complex dtype");
    return HPTTX_EINVAL;
  }
  s.dtype = it.dtype;
  s.A = it.in;
  s.B = it.out;
  s.rank = it.rank;
  for (uint32_t d = 0; d < it.rank; ++d) {
    s.in_size[d] = it.in_size[d];
    s.perm[d] = it.perm[d];
    s.in_outer[d] = it.in_outer[d];
    s.out_outer[d] = it.out_outer[d];
  }
  s.cc = it.complex_coef != 0 && cplx_type;
  if (it.dtype == HPTTX_F32 || it.dtype == HPTTX_C64)
    s.coef = CoefD{(double)(float)it.alpha_re, (double)(float)it.alpha_im,
                   (double)(float)it.beta_re, (double)(float)it.beta_im};
  else
    s.coef = CoefD{it.alpha_re, it.alpha_im, it.beta_re, it.beta_im};
  if (!s.cc) {
    s.coef.ai = 0.0;
    s.coef.bi = 0.0;
  }
  return HPTTX_OK;
}
}  // namespace hpttx

extern "C" {

int hpttx_batch_create(uint32_t count, const hpttx_batch_item *items, double tuning_timeout,
                       void *stream, hpttx_batch **out) {
  return hpttx_batch_create_ex(count, items, tuning_timeout, stream, 0, out);
}

int hpttx_batch_create_ex(uint32_t count, const hpttx_batch_item *items, double tuning_timeout,
                          void *stream, uint32_t flags, hpttx_batch **out) {
  if (!out) { hpttx::set_last_error("batch out-pointer is null"); return HPTTX_EINVAL; }
  if (flags & ~(uint32_t)HPTTX_BATCH_ONE_LAUNCH) {
    hpttx::set_last_error("unknown batch flags");
    return HPTTX_EINVAL;
  }
  *out = nullptr;
  if (count && !items) { hpttx::set_last_error("null item array"); return HPTTX_EINVAL; }
  std::vector<BatchItemSpec> specs(count);
  for (uint32_t k = 0; k < count; ++k) {
    int rc = batch_item_to_spec(items[k], k, specs[k], false);
    if (rc != HPTTX_OK) return rc;
  }
  hpttx_batch *h = new (std::nothrow) hpttx_batch();
  if (!h) { hpttx::set_last_error("out of host memory"); return HPTTX_ENOMEM; }
  std::string err;
  int rc = batch_build(specs, static_cast<hipStream_t>(stream), tuning_timeout,
                       (flags & HPTTX_BATCH_ONE_LAUNCH) ? kBatchOneLaunch : 0u, h->b, err);
  if (rc != kOk) {
    delete h;
    hpttx::set_last_error(err);
    return rc;
  }
  *out = h;
  return HPTTX_OK;
}

int hpttx_batch_exec_on(hpttx_batch *h, void *stream) {
  if (!h) { hpttx::set_last_error("null batch"); return HPTTX_EINVAL; }
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e;
  if (h->b.use_graph) {
    if (!h->b.gexec) {
      e = batch_capture(h->b);
      if (e != hipSuccess) {
        hpttx::set_last_error(std::string("graph capture: ") + hipGetErrorString(e));
        return HPTTX_EHIP;
      }
    }
    e = hipGraphLaunch(h->b.gexec, st);
  } else {
    e = batch_enqueue(h->b, st);
  }
  if (e == hipSuccess) e = hipEventRecord(h->b.last_exec, st);
  if (e != hipSuccess) {
    hpttx::set_last_error(std::string("batch launch: ") + hipGetErrorString(e));
    return HPTTX_EHIP;
  }
  return HPTTX_OK;
}

int hpttx_batch_exec(hpttx_batch *h) {
  if (!h) { hpttx::set_last_error("null batch"); return HPTTX_EINVAL; }
  return hpttx_batch_exec_on(h, h->b.stream);
}

int hpttx_batch_sync(hpttx_batch *h) {
  if (!h) { hpttx::set_last_error("null batch"); return HPTTX_EINVAL; }
  hipError_t e = hipStreamSynchronize(h->b.stream);
  if (e != hipSuccess) {
    hpttx::set_last_error(std::string("stream sync: ") + hipGetErrorString(e));
    return HPTTX_EHIP;
  }
  return HPTTX_OK;
}

int hpttx_batch_set_data(hpttx_batch *h, uint32_t index, const void *in, void *out) {
  if (!h) { hpttx::set_last_error("null batch"); return HPTTX_EINVAL; }
  std::string err;
  int rc = batch_set_data(h->b, index, in, out, err);
  if (rc != kOk) hpttx::set_last_error(err);
  return rc;
}

int hpttx_batch_use_graph(hpttx_batch *h, int enable) {
  if (!h) { hpttx::set_last_error("null batch"); return HPTTX_EINVAL; }
  h->b.use_graph = enable != 0;
  return HPTTX_OK;
}

int hpttx_batch_get_info(const hpttx_batch *h, hpttx_batch_info *info) {
  if (!h || !info) { hpttx::set_last_error("null argument"); return HPTTX_EINVAL; }
  std::memset(info, 0, sizeof(*info));
  info->count = h->b.count;
  info->groups = (uint32_t)h->b.groups.size();
  info->algorithmic_bytes = h->b.bytes;
  for (size_t i = 0; i < h->b.groups.size(); ++i) {
    info->tiles += h->b.groups[i].total_tiles;
    if (i < 4) {
      info->grid[i] = h->b.groups[i].grid;
      info->lds_bytes[i] = h->b.groups[i].lds_bytes;
    }
  }
  return HPTTX_OK;
}

int hpttx_batch_get_item_info(const hpttx_batch *h, uint32_t index, hpttx_batch_item_info *info) {
  if (!h || !info || index >= h->b.item_group.size()) {
    hpttx::set_last_error("bad batch, index or info pointer");
    return HPTTX_EINVAL;
  }
  const BatchGroup &g = h->b.groups[h->b.item_group[index]];
  const BatchEntry &e = g.h_ent[h->b.item_pos[index]];
  std::memset(info, 0, sizeof(*info));
  info->group = h->b.item_group[index];
  info->kernel = g.variant;
  info->variant = e.variant;
  info->lds_bytes = e.lds_bytes;
  info->tiles = e.ntiles;
  return HPTTX_OK;
}

int hpttx_batch_item_path(const hpttx_batch *h, uint32_t index, char *buf, size_t len) {
  if (!h || index >= h->b.paths.size()) { hpttx::set_last_error("bad batch or index"); return HPTTX_EINVAL; }
  const std::string &s = h->b.paths[index];
  if (buf && len) {
    size_t n = s.size() < len - 1 ? s.size() : len - 1;
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  return HPTTX_OK;
}

void hpttx_batch_destroy(hpttx_batch *h) {
  if (!h) return;
  // the last exec may have run on another stream than the batch's
  if (h->b.last_exec) (void)hipEventSynchronize(h->b.last_exec);
  (void)hipStreamSynchronize(h->b.stream);
  batch_free(h->b);
  delete h;
}

}  // extern "C"
