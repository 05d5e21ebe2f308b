// noalloc.hip -- hpttx_batch_run_noalloc: run a batch of transpositions
// without allocating device memory.
//
// hpttx_batch_create allocates the batch's device tables (entries, tile
// lookup, queue counters) with hipMalloc.  Callers that cannot allocate
// (code under stream capture, tight serving loops, harnesses that forbid
// hipMalloc) get this one-shot variant instead: the items are planned on the
// host, their tables are written into a static device array of fixed size
// (one copy per GPU, held until the process ends) and the batch runs as one
// persistent launch per element type over a dynamic work queue (the
// one-launch layout of hpttx_batch_create_ex, same kernels).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "batch.h"
#include "device.h"
#include "hptt_amd.h"

namespace hpttx {

void set_last_error(const std::string &m);
int batch_item_to_spec(const hpttx_batch_item &it, uint32_t k, BatchItemSpec &s,
                       bool allow_null_out);

namespace {

constexpr size_t kScratchBytes = HPTTX_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

// The static scratch.  In host code this name is only a handle for
// hipGetSymbolAddress: every device holds its own copy at its own address,
// and only that address (scratch_on_device) is passed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch[kMaxDevices];   // device address of `scratch`, per device

// Address of the current device's copy of `scratch` (looked up once per
// device, with that device current).
int scratch_on_device(void **out, std::string &err) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    err = std::string("no HIP device: ") + hipGetErrorString(e);
    return kEHip;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    err = "device index beyond the library's limit";
    return kEUnsupported;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      err = std::string("scratch address: ") + hipGetErrorString(e);
      return kEHip;
    }
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return kOk;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct NaGroup {
  int dtype = 0;
  std::vector<size_t> items;       // user indices, largest first
  std::vector<BatchEntry> ent;
  std::vector<uint32_t> begins;
  std::vector<uint32_t> chunks;
  uint32_t total = 0, lds = 0, grid = 0;
  size_t off_counter = 0, off_ent = 0, off_begins = 0, off_chunks = 0;
};

}  // namespace
}  // namespace hpttx

using namespace hpttx;

extern "C" int hpttx_batch_run_noalloc(uint32_t count, const hpttx_batch_item *items,
                                       void *stream) {
  if (count == 0) return HPTTX_OK;
  if (!items) {
    set_last_error("null item array");
    return HPTTX_EINVAL;
  }
  std::vector<BatchItemSpec> specs(count);
  for (uint32_t k = 0; k < count; ++k) {
    const int rc = batch_item_to_spec(items[k], k, specs[k], false);
    if (rc != HPTTX_OK) return rc;
  }
  std::string err;
  std::vector<PlannedItem> pl(count);
  for (uint32_t k = 0; k < count; ++k) {
    const int rc = plan_item(specs[k], TuneParams{}, pl[k], err);
    if (rc != kOk) {
      set_last_error("item " + std::to_string(k) + ": " + err);
      return rc;
    }
  }
  int rc = require_target_device(err);
  if (rc != kOk) {
    set_last_error(err);
    return rc;
  }
  // One group per element type (the kernel instance that dispatches each
  // entry's variant), largest transposition first.
  std::vector<NaGroup> groups;
  for (int dt = kF32; dt <= kC128; ++dt) {
    NaGroup g;
    g.dtype = dt;
    for (uint32_t k = 0; k < count; ++k)
      if (specs[k].dtype == dt) g.items.push_back(k);
    if (g.items.empty()) continue;
    std::stable_sort(g.items.begin(), g.items.end(),
                     [&](size_t x, size_t y) { return pl[x].p.N > pl[y].p.N; });
    uint64_t tiles = 0;
    for (size_t k : g.items) {
      BatchEntry e;
      make_entry(pl[k], specs[k], e);
      e.tile_begin = (uint32_t)tiles;
      g.begins.push_back((uint32_t)tiles);
      tiles += e.ntiles;
      g.lds = std::max(g.lds, batch_lds_bytes(e.lds_bytes));
      g.ent.push_back(e);
    }
    if (tiles > 0xFFFFFFFFull) {
      set_last_error("batch has more than 2^32 tiles of one element type");
      return HPTTX_EUNSUPPORTED;
    }
    g.total = (uint32_t)tiles;
    g.grid = batch_grid(dt, batch_group_key(0, true), g.lds, g.total, true);
    g.chunks = batch_queue_chunks(g.total, g.grid);
    groups.push_back(std::move(g));
  }
  // Scratch layout: per group a counter, the entries, their first tiles and
  // the queue's chunk boundaries, each 256-byte aligned.
  size_t off = 0;
  for (NaGroup &g : groups) {
    g.off_counter = off;
    off = align256(off + sizeof(uint32_t));
    g.off_ent = off;
    off = align256(off + sizeof(BatchEntry) * g.ent.size());
    g.off_begins = off;
    off = align256(off + sizeof(uint32_t) * g.begins.size());
    g.off_chunks = off;
    off = align256(off + sizeof(uint32_t) * g.chunks.size());
  }
  if (off > kScratchBytes) {
    set_last_error("batch tables need " + std::to_string(off) + " bytes; the static scratch of " +
                   "hpttx_batch_run_noalloc holds " + std::to_string(kScratchBytes) +
                   " (use hpttx_batch_create for larger batches)");
    return HPTTX_EINVAL;
  }
  void *dscratch = nullptr;
  rc = scratch_on_device(&dscratch, err);
  if (rc != kOk) {
    set_last_error(err);
    return rc;
  }
  // Host staging of the whole table image (counters zero).
  std::vector<unsigned char> image(off, 0);
  for (const NaGroup &g : groups) {
    std::memcpy(image.data() + g.off_ent, g.ent.data(), sizeof(BatchEntry) * g.ent.size());
    std::memcpy(image.data() + g.off_begins, g.begins.data(), sizeof(uint32_t) * g.begins.size());
    std::memcpy(image.data() + g.off_chunks, g.chunks.data(), sizeof(uint32_t) * g.chunks.size());
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  unsigned char *base = static_cast<unsigned char *>(dscratch);
  // The upload is ordered on the caller's stream; waiting for it keeps the
  // host image alive until the copy has read it.
  hipError_t e = hipMemcpyAsync(base, image.data(), off, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  for (size_t i = 0; i < groups.size() && e == hipSuccess; ++i) {
    const NaGroup &g = groups[i];
    BatchQueue q;
    q.counter = reinterpret_cast<uint32_t *>(base + g.off_counter);
    q.chunks = reinterpret_cast<const uint32_t *>(base + g.off_chunks);
    q.nchunks = (uint32_t)g.chunks.size() - 1;
    e = launch_batch_group(g.dtype, batch_group_key(0, true),
                           reinterpret_cast<const BatchEntry *>(base + g.off_ent),
                           reinterpret_cast<const uint32_t *>(base + g.off_begins), nullptr,
                           (uint32_t)g.ent.size(), g.total, g.lds, g.grid, st, &q);
  }
  if (e != hipSuccess) {
    set_last_error(std::string("noalloc batch: ") + hipGetErrorString(e));
    return HPTTX_EHIP;
  }
  return HPTTX_OK;
}
