// batch.h -- many independent transpositions in one persistent launch.
//
// No counterpart in the reference (one plan per transposition, run one after
// another: src/hptt.cc:16-93 + cgraph_trans.tcc:27-48).  Each transposition
// is planned like a single plan (planner.h); its block tiles are numbered
// consecutively, largest transposition first, and a persistent kernel per
// element type takes tiles from an atomic work queue.  The launches are
// captured in a hipGraph once and replayed.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "kernels.h"
#include "planner.h"

namespace hpttx {

// Device-side record of one transposition in a batch.
struct BatchEntry {
  uint32_t variant;      // bit0 rows, bit1 full read vec, bit2 full write vec,
                         // bit3 update, bit4 shared group, bit5 complex coef
  uint32_t tile_begin;   // first global tile index
  uint32_t ntiles;
  uint32_t lds_bytes;
  const void *A;
  void *B;
  double ar, ai, br, bi;
  union {
    TileDesc tile;
    RowDesc rows;
  } u;
};

struct BatchItemSpec {
  int dtype;
  const void *A;
  void *B;
  uint32_t rank;
  uint32_t in_size[kMaxRank];
  uint32_t perm[kMaxRank];
  uint32_t in_outer[kMaxRank];   // all 0 = unpadded
  uint32_t out_outer[kMaxRank];
  CoefD coef;
  bool cc;
};

// One element type's share of a batch.
struct BatchGroup {
  int dtype = 0;
  uint32_t nent = 0;
  uint32_t total_tiles = 0;
  uint32_t lds_bytes = 0;
  uint32_t grid = 0;
  BatchEntry *d_ent = nullptr;      // device table
  uint32_t *d_tile2ent = nullptr;   // device: tile -> entry
  std::vector<BatchEntry> h_ent;    // host copy
  std::vector<uint32_t> order;      // entry k <- user item order[k]
  uint64_t bytes = 0;               // algorithmic bytes
};

struct Batch {
  std::vector<BatchGroup> groups;
  uint32_t count = 0;
  uint32_t *d_counters = nullptr;   // one work counter per group
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool use_graph = true;
  uint64_t bytes = 0;
  std::vector<std::string> paths;   // per user item
};

// Plan every item (class tune params may be given per dtype/rank/path).
int batch_build(const std::vector<BatchItemSpec> &items, hipStream_t stream, double tuning_timeout,
                Batch &b, std::string &err);
// Enqueue: counters reset + one persistent kernel per group.
hipError_t batch_enqueue(Batch &b, hipStream_t stream);
// Capture batch_enqueue into a graph (done on first exec).
hipError_t batch_capture(Batch &b);
void batch_free(Batch &b);
// Re-point item k (user order) at new buffers; uploads the table entry.
int batch_set_data(Batch &b, uint32_t k, const void *A, void *B, std::string &err);

}  // namespace hpttx
