// batch.hip -- host side of batched transpositions: planning the items,
// building the launch groups and their device tables, launch grids,
// launches over side streams and the captured hipGraph.  The kernels are in
// batch_kernels.h (instantiated by batch_kernels_<type>.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>

#include "batch.h"
#include "batch_kernels.h"
#include "device.h"
#include "hpttx_core.h"
#include "tune.h"

namespace hpttx {

namespace {

struct KernelTable {
  const void *k[4][kVarCodes + 1];   // [kVarCodes]: the merged instance
  KernelTable() {
    batch_kernels_f32(k[kF32]);
    batch_kernels_f64(k[kF64]);
    batch_kernels_c64(k[kC64]);
    batch_kernels_c128(k[kC128]);
  }
};

const void *batch_kernel(int dtype, uint32_t variant) {
  static const KernelTable table;
  if (dtype < kF32 || dtype > kC128 || variant > (uint32_t)kVarCodes) return nullptr;
  return table.k[dtype][variant];
}

uint32_t entry_variant(const LaunchPlan &lp, bool update, bool cc, uint32_t vcap) {
  uint32_t v = 0;
  if (lp.kind == kRows) {
    v |= kVarRows;
    if (lp.rows.v == vcap && vcap > 1) v |= lp.rows.ua ? kVarVw : kVarVr;
  } else {
    if (vcap > 1 && !lp.tile.ua_r) v |= kVarVr;
    if (vcap > 1 && !lp.tile.ua_w) v |= kVarVw;
    if (lp.tile.VS > 1) v |= kVarSh;
  }
  if (update) v |= kVarUpd;
  if (cc) v |= kVarCc;
  // the batch kernels stream loads and stores alike, or neither
  if (lp.nt_load && lp.nt_store) v |= kVarNt;
  return v;
}

// Finalize vectors for the pointers (full 16-byte vectors or scalars).
void pick_vectors(const Problem &p, LaunchPlan &lp, const void *A, const void *B) {
  const uint32_t cap = 16 / p.esize;
  uint32_t vr = lp.vr_shape, vw = lp.vw_shape;
  if (lp.kind == kRows) {
    uint32_t v = vr;
    if (v > align_vec(A, p.esize, cap) || v > align_vec(B, p.esize, cap) || v != cap) v = 1;
    vr = vw = v;
  } else {
    if (vr > align_vec(A, p.esize, cap) || vr != cap) vr = 1;
    if (vw > align_vec(B, p.esize, cap) || vw != cap) vw = 1;
  }
  finalize_vectors(p, lp, vr, vw);
  stream_for_pointers(p, lp, A, B);
}

}  // namespace

int plan_item(const BatchItemSpec &it, const TuneParams &tp, PlannedItem &q, std::string &err) {
  if (!it.A || !it.B) { err = "null data pointer"; return kEInval; }
  bool padded_in = false, padded_out = false;
  for (uint32_t d = 0; d < it.rank && d < (uint32_t)kMaxRank; ++d) {
    padded_in |= it.in_outer[d] != 0;
    padded_out |= it.out_outer[d] != 0;
  }
  int rc = build_problem(it.dtype, it.rank, it.in_size, it.perm, padded_in ? it.in_outer : nullptr,
                         padded_out ? it.out_outer : nullptr, q.p, err);
  if (rc != kOk) return rc;
  q.cc = it.cc && (it.coef.ai != 0.0 || it.coef.bi != 0.0);
  q.update = q.cc ? (std::fabs(it.coef.br) >= 1e-16 || std::fabs(it.coef.bi) >= 1e-16)
                  : std::fabs(it.coef.br) >= 1e-16;
  rc = plan_launch(q.p, tp, q.lp, err);
  if (rc != kOk) return rc;
  pick_vectors(q.p, q.lp, it.A, it.B);
  return kOk;
}

std::string class_key(const PlannedItem &q) {
  std::ostringstream os;
  int bucket = 0;
  for (uint64_t n = q.p.N; n > 1; n >>= 2) ++bucket;
  os << q.p.dtype << '|' << q.p.fr << '|' << q.lp.path << '|' << bucket << '|' << q.update;
  return os.str();
}

void make_entry(const PlannedItem &q, const BatchItemSpec &it, BatchEntry &e) {
  std::memset(&e, 0, sizeof(e));
  e.variant = entry_variant(q.lp, q.update, q.cc, 16 / q.p.esize);
  e.ntiles = q.lp.kind == kTile ? q.lp.tile.ntiles : q.lp.rows.ntiles;
  e.lds_bytes = (uint32_t)q.lp.lds_bytes;
  e.A = it.A;
  e.B = it.B;
  e.ar = it.coef.ar;
  e.ai = q.cc ? it.coef.ai : 0.0;
  e.br = it.coef.br;
  e.bi = q.cc ? it.coef.bi : 0.0;
  if (q.lp.kind == kTile) e.u.tile = q.lp.tile;
  else e.u.rows = q.lp.rows;
}

uint32_t batch_lds_bytes(uint32_t lds_bytes) { return (lds_bytes + 15u) & ~15u; }

uint32_t batch_lds_class(uint32_t lds_bytes) { return (lds_bytes + 8191u) / 8192u; }

// Clears a dynamic queue's counter (a kernel node, so that a captured graph
// replays it; atomic, so the next launch's atomic increments on every XCD
// see it).
__global__ void k_queue_reset(uint32_t *counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) atomicExch(counter, 0u);
}

std::vector<uint32_t> batch_queue_chunks(uint32_t total, uint32_t grid) {
  std::vector<uint32_t> c;
  const uint64_t div = 4ull * std::max<uint32_t>(grid, 1);
  uint32_t t = 0;
  while (t < total) {
    c.push_back(t);
    const uint32_t rem = total - t;
    const uint32_t n = (uint32_t)std::max<uint64_t>(1, rem / div);
    t += std::min(n, rem);
  }
  c.push_back(total);
  return c;
}

bool batch_persistent() {
  const char *m = std::getenv("HPTTX_BATCH_GRID");
  return m && std::strcmp(m, "persistent") == 0;
}

uint32_t batch_grid(int dtype, uint32_t variant, uint32_t lds_bytes, uint32_t total_tiles,
                    bool persistent) {
  total_tiles = std::max<uint32_t>(1, total_tiles);
  if (!persistent) return total_tiles;
  static std::mutex mu;
  static std::map<std::tuple<int, uint32_t, uint32_t>, uint32_t> cache;
  static int ncu = 0;
  uint32_t per_gpu;
  {
    std::lock_guard<std::mutex> lk(mu);
    if (!ncu) {
      int dev = 0;
      (void)hipGetDevice(&dev);
      if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
          ncu < 1)
        ncu = 256;
    }
    auto key = std::make_tuple(dtype, variant, lds_bytes);
    auto it = cache.find(key);
    if (it != cache.end()) {
      per_gpu = it->second;
    } else {
      int occ = 1;
      const void *k = batch_kernel(dtype, variant);
      if (!k || hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, kBlock, lds_bytes) != hipSuccess ||
          occ < 1) {
        (void)hipGetLastError();
        occ = 1;
      }
      per_gpu = (uint32_t)(ncu * occ);
      cache[key] = per_gpu;
    }
  }
  return std::min<uint32_t>(per_gpu, total_tiles);
}

std::vector<uint32_t> batch_entry_of_tile(const std::vector<uint32_t> &begins, uint32_t total) {
  std::vector<uint32_t> t(total, 0);
  const uint32_t nent = (uint32_t)begins.size();
  for (uint32_t e = 0; e < nent; ++e) {
    const uint32_t end = e + 1 < nent ? begins[e + 1] : total;
    for (uint32_t k = begins[e]; k < end; ++k) t[k] = e;
  }
  return t;
}

hipError_t launch_batch_group(int dtype, uint32_t variant, const BatchEntry *d_ent,
                              const uint32_t *d_begins, const uint32_t *d_ent_of, uint32_t nent,
                              uint32_t total, uint32_t lds_bytes, uint32_t grid, hipStream_t st,
                              const BatchQueue *queue) {
  if (!total || !nent) return hipSuccess;
  const void *k = batch_kernel(dtype, variant);
  if (!k) return hipErrorInvalidValue;
  uint32_t *counter = queue ? queue->counter : nullptr;
  const uint32_t *chunks = queue ? queue->chunks : nullptr;
  uint32_t nchunks = queue ? queue->nchunks : 0;
  if (counter) {
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, counter);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  void *args[] = {(void *)&d_ent, (void *)&d_begins, (void *)&d_ent_of, (void *)&nent,
                  (void *)&total, (void *)&counter, (void *)&chunks, (void *)&nchunks};
  // 2-D grid: x * 256 threads must stay below 2^32 work items per dimension.
  const uint32_t gx = grid < kGridX ? grid : kGridX;
  return hipLaunchKernel(k, dim3(gx, (grid + gx - 1) / gx), dim3(kBlock), args, lds_bytes, st);
}

uint32_t batch_group_key(uint32_t variant, bool merged) { return merged ? (uint32_t)kVarCodes : variant; }

bool batch_merged_default(bool fallback) {
  const char *m = std::getenv("HPTTX_BATCH_MODE");
  if (m && std::string(m) == "split") return false;
  if (m && std::string(m) == "merged") return true;
  return fallback;
}

namespace {

// Scratch a class tuning run may use (A and B copies together).
constexpr uint64_t kTuneScratchBytes = 256ull << 20;

uint64_t round256(uint64_t x) { return (x + 255) / 256 * 256; }

// Batch-context tuning of one plan class: K copies of the class's largest
// member, on scratch buffers, run as one batch launch per candidate tile
// shape, and the fastest shape wins.  A lone member would be timed on its
// latency (favouring small tiles); a batch is bound by throughput.
int tune_class_batch(const PlannedItem &rep, const BatchItemSpec &spec, double budget_s,
                     hipStream_t st, TuneParams &out, std::string &err) {
  out = TuneParams{};   // the model's pick unless a candidate beats it
  const uint64_t es = rep.p.esize;
  const uint64_t a_bytes = round256(rep.p.spanA * es), b_bytes = round256(rep.p.spanB * es);
  const uint64_t K = std::max<uint64_t>(1, std::min<uint64_t>(64, kTuneScratchBytes / (a_bytes + b_bytes)));
  char *A = nullptr, *B = nullptr;
  if (hipMalloc(&A, K * a_bytes) != hipSuccess || hipMalloc(&B, K * b_bytes) != hipSuccess) {
    (void)hipGetLastError();
    if (A) (void)hipFree(A);
    return kOk;   // no room for scratch: keep the model pick
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  BatchEntry *d_ent = nullptr;
  uint32_t *d_begins = nullptr, *d_ent_of = nullptr;
  size_t ent_of_cap = 0;
  int rc = kOk;
  // Every scratch copy starts as the member's own A and B (read only): timings
  // on zero-filled buffers favour other shapes than real data does.
  bool filled = true;
  for (uint64_t k = 0; k < K && filled; ++k)
    filled = hipMemcpyAsync(A + k * a_bytes, spec.A, rep.p.spanA * es, hipMemcpyDeviceToDevice,
                            st) == hipSuccess &&
             hipMemcpyAsync(B + k * b_bytes, spec.B, rep.p.spanB * es, hipMemcpyDeviceToDevice,
                            st) == hipSuccess;
  if (!filled || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
      hipMalloc(&d_ent, sizeof(BatchEntry) * K) != hipSuccess ||
      hipMalloc(&d_begins, sizeof(uint32_t) * K) != hipSuccess) {
    err = "batch tuning: HIP setup failed";
    rc = kEHip;
  }
  struct Cand {
    TuneParams tp;
    std::vector<BatchEntry> ent;
    std::vector<uint32_t> begins;
    uint32_t total = 0, lds = 0, key = 0;
  };
  // Build a candidate's K entries (tile shape from tp, pointers into scratch).
  auto build = [&](const TuneParams &tp, Cand &c) -> bool {
    BatchItemSpec s = spec;
    s.A = A;
    s.B = B;
    PlannedItem q;
    std::string e2;
    if (plan_item(s, tp, q, e2) != kOk) return false;
    c.tp = tp;
    c.ent.assign(K, BatchEntry{});
    c.begins.assign(K, 0);
    uint64_t total = 0;
    for (uint64_t k = 0; k < K; ++k) {
      s.A = A + k * a_bytes;
      s.B = B + k * b_bytes;
      make_entry(q, s, c.ent[k]);
      c.ent[k].tile_begin = (uint32_t)total;
      c.begins[k] = (uint32_t)total;
      total += c.ent[k].ntiles;
    }
    if (total > 0xFFFFFFFFull) return false;
    c.total = (uint32_t)total;
    c.lds = batch_lds_bytes(c.ent[0].lds_bytes);
    c.key = batch_group_key(c.ent[0].variant, false);
    return true;
  };
  // Upload a candidate's tables and time `reps` launches (ms per launch).
  auto time_cand = [&](const Cand &c, int reps, float &ms) -> bool {
    const std::vector<uint32_t> ent_of = batch_entry_of_tile(c.begins, c.total);
    if (ent_of.size() > ent_of_cap) {
      if (d_ent_of) (void)hipFree(d_ent_of);
      d_ent_of = nullptr;
      if (hipMalloc(&d_ent_of, sizeof(uint32_t) * ent_of.size()) != hipSuccess) return false;
      ent_of_cap = ent_of.size();
    }
    if (hipMemcpyAsync(d_ent, c.ent.data(), sizeof(BatchEntry) * K, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_begins, c.begins.data(), sizeof(uint32_t) * K, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_ent_of, ent_of.data(), sizeof(uint32_t) * ent_of.size(), hipMemcpyHostToDevice, st) != hipSuccess)
      return false;
    const uint32_t grid = batch_grid(spec.dtype, c.key, c.lds, c.total, batch_persistent());
    if (launch_batch_group(spec.dtype, c.key, d_ent, d_begins, d_ent_of, (uint32_t)K, c.total, c.lds, grid, st) != hipSuccess)
      return false;
    (void)hipEventRecord(e0, st);
    for (int i = 0; i < reps; ++i)
      (void)launch_batch_group(spec.dtype, c.key, d_ent, d_begins, d_ent_of, (uint32_t)K, c.total, c.lds, grid, st);
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess) return false;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms /= reps;
    return true;
  };
  const auto t_start = std::chrono::steady_clock::now();
  std::vector<std::pair<float, TuneParams>> scored;
  std::vector<LaunchPlan> seen;
  if (rc == kOk) {
    for (const TuneParams &tp : tune_candidates(rep.p)) {
      if (budget_s > 0 &&
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count() > budget_s)
        break;
      // plan once to skip candidates whose launch shape was already timed
      PlannedItem q;
      std::string e2;
      BatchItemSpec s = spec;
      s.A = A;
      s.B = B;
      if (plan_item(s, tp, q, e2) != kOk) continue;
      bool dup = false;
      for (const LaunchPlan &l : seen) dup = dup || same_launch(l, q.lp);
      if (dup) continue;
      seen.push_back(q.lp);
      Cand c;
      float ms = 0;
      if (!build(tp, c)) continue;
      if (!time_cand(c, 3, ms)) {
        err = "batch tuning: launch failed";
        rc = kEHip;
        break;
      }
      scored.push_back({ms, tp});
    }
  }
  if (rc == kOk && !scored.empty()) {
    std::sort(scored.begin(), scored.end(),
              [](const std::pair<float, TuneParams> &x, const std::pair<float, TuneParams> &y) {
                return x.first < y.first;
              });
    float best = 1e30f;
    for (size_t i = 0; i < scored.size() && i < 3; ++i) {
      Cand c;
      float ms = 0;
      if (!build(scored[i].second, c) || !time_cand(c, 10, ms)) continue;
      if (ms < best) {
        best = ms;
        out = scored[i].second;
      }
    }
  }
  (void)hipStreamSynchronize(st);
  (void)hipGetLastError();
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (d_ent) (void)hipFree(d_ent);
  if (d_begins) (void)hipFree(d_begins);
  if (d_ent_of) (void)hipFree(d_ent_of);
  (void)hipFree(A);
  (void)hipFree(B);
  return rc;
}

}  // namespace

namespace {

// Lay a batch out from its planned items: launch groups, device tables, side
// streams and events.
int batch_assemble(const std::vector<BatchItemSpec> &items, const std::vector<PlannedItem> &pl,
                   hipStream_t stream, uint32_t flags, Batch &b, std::string &err);

// Median whole-batch time (ms) of `reps` direct launches of the batch laid
// out from `pl` over `items` (items whose outputs are scratch buffers).
int time_layout(const std::vector<BatchItemSpec> &items, const std::vector<PlannedItem> &pl,
                hipStream_t st, uint32_t flags, int reps, float &ms, std::string &err) {
  Batch t;
  int rc = batch_assemble(items, pl, st, flags, t, err);
  if (rc != kOk) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<float> v;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = batch_enqueue(t, st);   // warm-up
  for (int r = 0; r < reps && e == hipSuccess; ++r) {
    e = hipEventRecord(e0, st);
    if (e == hipSuccess) e = batch_enqueue(t, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float x = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&x, e0, e1);
    v.push_back(x);
  }
  (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  batch_free(t);
  if (e != hipSuccess) {
    err = std::string("batch tuning: ") + hipGetErrorString(e);
    return kEHip;
  }
  std::sort(v.begin(), v.end());
  ms = v[v.size() / 2];
  return kOk;
}

// Whole-batch selection of the classes' tuned tile shapes.  Per-class
// timing sees one class alone; in the batch the classes share the GPU, and
// a shape that wins alone can lose there (measured: per-class tuning alone
// made the 4096-tensor manifests 10% slower).  So the batch starts from the
// model's plans and takes each class's tuned plans, largest classes first,
// only where the whole batch (run on scratch copies of the outputs, each at
// its output's alignment) gets at least 1% faster.  The result is never
// slower than the model beyond timing noise.  Without room for the scratch
// outputs the model's plans are kept.
int select_whole_batch(const std::vector<BatchItemSpec> &items,
                       const std::vector<PlannedItem> &pl_model,
                       const std::vector<PlannedItem> &pl_tuned,
                       const std::vector<std::string> &keys, hipStream_t st, uint32_t flags,
                       double budget_s, std::vector<PlannedItem> &out, std::string &err) {
  out = pl_model;
  const auto t_start = std::chrono::steady_clock::now();
  // classes whose tuned plans differ from the model's, by bytes
  std::map<std::string, uint64_t> cls_bytes;
  for (size_t k = 0; k < items.size(); ++k)
    if (!same_launch(pl_model[k].lp, pl_tuned[k].lp))
      cls_bytes[keys[k]] += pl_model[k].p.N * pl_model[k].p.esize;
  if (cls_bytes.empty()) return kOk;
  std::vector<std::pair<uint64_t, std::string>> order;
  for (auto &kv : cls_bytes) order.push_back({kv.second, kv.first});
  std::sort(order.rbegin(), order.rend());
  // scratch outputs at each output's offset modulo 256 bytes
  std::vector<uint64_t> off(items.size());
  uint64_t total = 0;
  for (size_t k = 0; k < items.size(); ++k) {
    const uint64_t mis = (uint64_t)(uintptr_t)items[k].B % 256;
    off[k] = total + mis;
    total = round256(off[k] + pl_model[k].p.spanB * pl_model[k].p.esize);
  }
  char *scratch = nullptr;
  if (hipMalloc(&scratch, total) != hipSuccess) {
    (void)hipGetLastError();
    return kOk;   // no room: the model's plans
  }
  std::vector<BatchItemSpec> trial_items = items;
  hipError_t e = hipSuccess;
  for (size_t k = 0; k < items.size() && e == hipSuccess; ++k) {
    trial_items[k].B = scratch + off[k];
    e = hipMemcpyAsync(scratch + off[k], items[k].B, pl_model[k].p.spanB * pl_model[k].p.esize,
                       hipMemcpyDeviceToDevice, st);
  }
  int rc = kOk;
  float best = 0;
  if (e != hipSuccess) {
    err = std::string("batch tuning: ") + hipGetErrorString(e);
    rc = kEHip;
  } else {
    rc = time_layout(trial_items, out, st, flags, 5, best, err);
  }
  for (size_t c = 0; c < order.size() && rc == kOk; ++c) {
    if (budget_s > 0 &&
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count() > budget_s)
      break;
    std::vector<PlannedItem> trial = out;
    for (size_t k = 0; k < items.size(); ++k)
      if (keys[k] == order[c].second) trial[k] = pl_tuned[k];
    float ms = 0;
    rc = time_layout(trial_items, trial, st, flags, 5, ms, err);
    if (rc == kOk && ms < 0.99f * best) {
      best = ms;
      out = trial;
    }
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(scratch);
  if (rc != kOk) out = pl_model;
  return rc;
}

}  // namespace

int batch_build(const std::vector<BatchItemSpec> &items, hipStream_t stream, double tuning_timeout,
                uint32_t flags, Batch &b, std::string &err) {
  std::vector<PlannedItem> pl(items.size());
  std::vector<std::string> keys(items.size());
  std::map<std::string, size_t> class_rep;   // largest member of each class
  for (size_t k = 0; k < items.size(); ++k) {
    int rc = plan_item(items[k], TuneParams{}, pl[k], err);
    if (rc != kOk) { err = "item " + std::to_string(k) + ": " + err; return rc; }
    keys[k] = class_key(pl[k]);
    auto it2 = class_rep.find(keys[k]);
    if (it2 == class_rep.end() || pl[it2->second].p.N < pl[k].p.N) class_rep[keys[k]] = k;
  }
  // Items are validated first, then the device (as for single plans).
  {
    int rc = require_target_device(err);
    if (rc != kOk) return rc;
  }
  // Per-class tile shapes: tuned in batch context on copies of the class's
  // largest member (scratch buffers), or taken from the tuning cache (under
  // a batch key: a lone plan of the same problem may tune differently); then
  // kept class by class only where the whole batch gets faster.  Half of a
  // positive budget goes to each step.
  if (tuning_timeout != 0.0) {
    const auto t_start = std::chrono::steady_clock::now();
    const double share =
        tuning_timeout > 0 ? 0.5 * tuning_timeout / std::max<size_t>(1, class_rep.size()) : -1.0;
    std::map<std::string, TuneParams> class_tune;
    for (auto &kv : class_rep) {
      const PlannedItem &rep = pl[kv.second];
      const std::string key = "batch|" + tune_key(rep.p, rep.update, rep.cc);
      TuneParams tp{};
      if (!tune_lookup_key(key, tp)) {
        int rc = tune_class_batch(rep, items[kv.second], share, stream, tp, err);
        if (rc != kOk) return rc;
        tune_store_key(key, tp);
      }
      class_tune[kv.first] = tp;
    }
    // A class's tuned shape may not fit every member (a larger tile than
    // 64 KiB of LDS for a member with other extents): such members keep
    // their model plan.
    std::vector<PlannedItem> pl_tuned = pl;
    for (size_t k = 0; k < items.size(); ++k) {
      PlannedItem tuned;
      std::string e2;
      if (plan_item(items[k], class_tune[keys[k]], tuned, e2) == kOk) pl_tuned[k] = tuned;
    }
    double left = -1.0;
    if (tuning_timeout > 0)
      left = std::max(0.0, tuning_timeout - std::chrono::duration<double>(
                                                std::chrono::steady_clock::now() - t_start)
                                                .count());
    std::vector<PlannedItem> chosen;
    int rc = select_whole_batch(items, pl, pl_tuned, keys, stream, flags,
                                tuning_timeout > 0 ? std::max(left, 1e-3) : -1.0, chosen, err);
    if (rc != kOk) return rc;
    pl = chosen;
  }
  return batch_assemble(items, pl, stream, flags, b, err);
}

namespace {

int batch_assemble(const std::vector<BatchItemSpec> &items, const std::vector<PlannedItem> &pl,
                   hipStream_t stream, uint32_t flags, Batch &b, std::string &err) {
  b = Batch();
  b.count = (uint32_t)items.size();
  b.stream = stream;
  b.paths.resize(items.size());
  b.item_group.assign(items.size(), 0);
  b.item_pos.assign(items.size(), 0);
  // Batches: one launch per (element type, variant), forked over streams
  // (measured faster than one dispatching launch per type on the 4096-tensor
  // mix: 4.93 ms vs 5.53 ms fp32, 8.71 vs 10.18 ms fp64).
  // The one-launch layout: one persistent launch per element type over a
  // dynamic queue.  Otherwise the environment may still ask for merged
  // launches (HPTTX_BATCH_MODE=merged), persistent grids
  // (HPTTX_BATCH_GRID=persistent) or the dynamic queue (HPTTX_BATCH_GRID=
  // queue, persistent grids with a dynamic queue).
  const char *grid_env = std::getenv("HPTTX_BATCH_GRID");
  const bool env_queue = grid_env && std::strcmp(grid_env, "queue") == 0;
  const bool one_launch = (flags & kBatchOneLaunch) != 0;
  const bool merged = one_launch || batch_merged_default(false);
  const bool use_queue = one_launch || env_queue;
  const bool persistent = use_queue || batch_persistent();
  // Groups: (element type, kernel, LDS class).  A launch reserves its
  // largest entry's LDS for every block, so entries are also split by LDS
  // need (8 KiB classes): small tiles keep their occupancy.  Merged mode
  // keeps one launch per element type.
  std::map<std::tuple<int, uint32_t, uint32_t>, std::vector<size_t>> by_kernel;
  for (size_t k = 0; k < items.size(); ++k) {
    b.paths[k] = pl[k].lp.path;
    BatchEntry probe;
    make_entry(pl[k], items[k], probe);
    by_kernel[std::make_tuple(items[k].dtype, batch_group_key(probe.variant, merged),
                              merged ? 0u : batch_lds_class(probe.lds_bytes))]
        .push_back(k);
  }
  hipError_t he = hipSuccess;
  for (auto &kv : by_kernel) {
    BatchGroup g;
    g.dtype = std::get<0>(kv.first);
    g.variant = std::get<1>(kv.first);
    std::vector<size_t> idx = kv.second;
    // Largest transposition first: its tiles are handed out first and the
    // small ones fill the tail.
    std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) {
      return pl[x].p.N > pl[y].p.N;
    });
    uint64_t tiles = 0;
    std::vector<uint32_t> begins;
    for (size_t k : idx) {
      BatchEntry e;
      make_entry(pl[k], items[k], e);
      e.tile_begin = (uint32_t)tiles;
      begins.push_back((uint32_t)tiles);
      tiles += e.ntiles;
      g.lds_bytes = std::max(g.lds_bytes, e.lds_bytes);
      g.bytes += pl[k].p.N * pl[k].p.esize * (uint64_t)(pl[k].update ? 3 : 2);
      g.h_ent.push_back(e);
      g.order.push_back((uint32_t)k);
    }
    if (tiles > 0xFFFFFFFFull) { err = "batch has more than 2^32 tiles"; return kEUnsupported; }
    g.nent = (uint32_t)g.h_ent.size();
    g.total_tiles = (uint32_t)tiles;
    g.lds_bytes = batch_lds_bytes(g.lds_bytes);
    g.grid = batch_grid(g.dtype, g.variant, g.lds_bytes, g.total_tiles, persistent);
    const std::vector<uint32_t> ent_of = batch_entry_of_tile(begins, g.total_tiles);
    std::vector<uint32_t> chunks;
    if (use_queue) {
      chunks = batch_queue_chunks(g.total_tiles, g.grid);
      g.nchunks = (uint32_t)chunks.size() - 1;
      he = hipMalloc(&g.d_queue, sizeof(uint32_t));
      if (he == hipSuccess) he = hipMalloc(&g.d_chunks, sizeof(uint32_t) * chunks.size());
      if (he == hipSuccess)
        he = hipMemcpy(g.d_chunks, chunks.data(), sizeof(uint32_t) * chunks.size(),
                       hipMemcpyHostToDevice);
    }
    if (he == hipSuccess) he = hipMalloc(&g.d_ent, sizeof(BatchEntry) * g.nent);
    if (he == hipSuccess) he = hipMalloc(&g.d_begins, sizeof(uint32_t) * g.nent);
    if (he == hipSuccess) he = hipMalloc(&g.d_ent_of, sizeof(uint32_t) * ent_of.size());
    if (he == hipSuccess)
      he = hipMemcpy(g.d_ent, g.h_ent.data(), sizeof(BatchEntry) * g.nent, hipMemcpyHostToDevice);
    if (he == hipSuccess)
      he = hipMemcpy(g.d_begins, begins.data(), sizeof(uint32_t) * g.nent, hipMemcpyHostToDevice);
    if (he == hipSuccess)
      he = hipMemcpy(g.d_ent_of, ent_of.data(), sizeof(uint32_t) * ent_of.size(),
                     hipMemcpyHostToDevice);
    b.bytes += g.bytes;
    for (uint32_t i = 0; i < (uint32_t)g.order.size(); ++i) {
      b.item_group[g.order[i]] = (uint32_t)b.groups.size();
      b.item_pos[g.order[i]] = i;
    }
    b.groups.push_back(std::move(g));
    if (he != hipSuccess) break;
  }
  // Side streams for concurrent groups; groups go, largest first, to the
  // least loaded lane.
  const size_t nside = std::min<size_t>(b.groups.size() > 0 ? b.groups.size() - 1 : 0, kBatchLanes - 1);
  for (size_t i = 0; i < nside && he == hipSuccess; ++i) {
    hipStream_t s = nullptr;
    he = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (he == hipSuccess) b.side.push_back(s);
  }
  for (size_t i = 0; i < nside + 1 && he == hipSuccess; ++i) {
    hipEvent_t e = nullptr;
    he = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (he == hipSuccess) b.events.push_back(e);
  }
  if (he == hipSuccess) he = hipEventCreateWithFlags(&b.last_exec, hipEventDisableTiming);
  if (he == hipSuccess) {
    std::vector<size_t> gorder(b.groups.size());
    for (size_t i = 0; i < gorder.size(); ++i) gorder[i] = i;
    std::sort(gorder.begin(), gorder.end(),
              [&](size_t x, size_t y) { return b.groups[x].bytes > b.groups[y].bytes; });
    std::vector<uint64_t> load(nside + 1, 0);
    b.lane.assign(b.groups.size(), 0);
    for (size_t gi : gorder) {
      size_t best = 0;
      for (size_t l = 1; l < load.size(); ++l)
        if (load[l] < load[best]) best = l;
      b.lane[gi] = (int)best;
      load[best] += b.groups[gi].bytes;
    }
  }
  if (he != hipSuccess) {
    err = std::string("batch setup: ") + hipGetErrorString(he);
    batch_free(b);
    return he == hipErrorOutOfMemory ? kENoMem : kEHip;
  }
  return kOk;
}

}  // namespace

hipError_t batch_enqueue(Batch &b, hipStream_t st) {
  if (b.groups.empty()) return hipSuccess;
  hipError_t e = hipSuccess;
  if (!b.side.empty()) e = hipEventRecord(b.events[0], st);
  for (size_t i = 0; i < b.side.size() && e == hipSuccess; ++i)
    e = hipStreamWaitEvent(b.side[i], b.events[0], 0);
  for (size_t i = 0; i < b.groups.size() && e == hipSuccess; ++i) {
    const BatchGroup &g = b.groups[i];
    hipStream_t ls = b.lane[i] == 0 ? st : b.side[b.lane[i] - 1];
    BatchQueue q;
    q.counter = g.d_queue;
    q.chunks = g.d_chunks;
    q.nchunks = g.nchunks;
    e = launch_batch_group(g.dtype, g.variant, g.d_ent, g.d_begins, g.d_ent_of, g.nent,
                           g.total_tiles, g.lds_bytes, g.grid, ls, g.d_queue ? &q : nullptr);
  }
  for (size_t i = 0; i < b.side.size() && e == hipSuccess; ++i) {
    e = hipEventRecord(b.events[1 + i], b.side[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(st, b.events[1 + i], 0);
  }
  return e;
}

hipError_t batch_capture(Batch &b) {
  hipStream_t cs = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) return e;
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    hipError_t e2 = batch_enqueue(b, cs);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    if (e2 != hipSuccess) e = e2;
    if (e == hipSuccess) {
      b.graph = g;
      e = hipGraphInstantiate(&b.gexec, g, nullptr, nullptr, 0);
    } else if (g) {
      (void)hipGraphDestroy(g);
    }
  }
  (void)hipStreamDestroy(cs);
  return e;
}

void batch_free(Batch &b) {
  if (b.gexec) (void)hipGraphExecDestroy(b.gexec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  for (BatchGroup &g : b.groups) {
    if (g.d_ent) (void)hipFree(g.d_ent);
    if (g.d_begins) (void)hipFree(g.d_begins);
    if (g.d_ent_of) (void)hipFree(g.d_ent_of);
    if (g.d_queue) (void)hipFree(g.d_queue);
    if (g.d_chunks) (void)hipFree(g.d_chunks);
    g.d_ent_of = nullptr;
    g.d_queue = nullptr;
    g.d_chunks = nullptr;
    g.d_ent = nullptr;
    g.d_begins = nullptr;
  }
  for (hipStream_t s : b.side) (void)hipStreamDestroy(s);
  for (hipEvent_t e : b.events) (void)hipEventDestroy(e);
  if (b.last_exec) (void)hipEventDestroy(b.last_exec);
  b.last_exec = nullptr;
  b.side.clear();
  b.events.clear();
  b.gexec = nullptr;
  b.graph = nullptr;
  b.groups.clear();
}

int batch_set_data(Batch &b, uint32_t k, const void *A, void *B, std::string &err) {
  if (!A || !B) { err = "null data pointer"; return kEInval; }
  if (k >= b.item_group.size()) { err = "item index out of range"; return kEInval; }
  BatchGroup &g = b.groups[b.item_group[k]];
  const uint32_t i = b.item_pos[k];
  BatchEntry &e = g.h_ent[i];
  const uint32_t es = dtype_esize(g.dtype);
  // Exact-vector sides were chosen for the old pointers' 16-byte alignment;
  // new pointers must keep it there (unaligned sides accept any
  // element-aligned pointer).
  const bool vr = (e.variant & kVarVr) != 0, vw = (e.variant & kVarVw) != 0;
  const bool rows = (e.variant & kVarRows) != 0;
  const bool need_a = vr, need_b = rows ? vr : vw;
  const uint32_t cap = 16 / es;
  if ((need_a && align_vec(A, es, cap) != cap) || (need_b && align_vec(B, es, cap) != cap)) {
    err = "new pointers are not 16-byte aligned like the ones the batch was built for";
    return kEInval;
  }
  // Execs still running (on any stream) read the table: wait for the last
  // one, then rewrite the entry synchronously.
  hipError_t he = b.last_exec ? hipEventSynchronize(b.last_exec) : hipSuccess;
  if (he == hipSuccess) {
    e.A = A;
    e.B = B;
    he = hipMemcpy(&g.d_ent[i], &e, sizeof(BatchEntry), hipMemcpyHostToDevice);
  }
  if (he != hipSuccess) { err = hipGetErrorString(he); return kEHip; }
  return kOk;
}

}  // namespace hpttx
