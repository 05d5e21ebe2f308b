// batch.h -- many independent transpositions in a few graph-captured launches.
//
// No counterpart in the reference (one plan per transposition, run one after
// another: src/hptt.cc:16-93 + cgraph_trans.tcc:27-48).  Each transposition
// is planned like a single plan (planner.h); its block tiles are numbered
// consecutively, largest transposition first, and one kernel launch per
// group runs one block per tile (in an XCD-aware order), finding each
// tile's transposition by a binary search over the entries' first tiles.
// The launches are captured in a hipGraph once and replayed.  The serving
// loop (serve.h) drives the same kernel with per-launch tables.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "kernels.h"
#include "planner.h"

namespace hpttx {

// Streams a batch spreads its launches over (the exec stream plus side ones).
constexpr size_t kBatchLanes = 4;

// Device-side record of one transposition in a batch.
struct BatchEntry {
  uint32_t variant;      // bit0 rows, bit1 full read vec, bit2 full write vec,
                         // bit3 update, bit4 shared group, bit5 complex coef,
                         // bit6 streaming (non-temporal) loads and stores
  uint32_t tile_begin;   // first tile of this entry within its launch
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

// One item planned: problem, launch plan (vectors fixed for its pointers),
// update flag, effective complex-coefficient flag.
struct PlannedItem {
  Problem p;
  LaunchPlan lp;
  bool update = false;
  bool cc = false;
};

int plan_item(const BatchItemSpec &it, const TuneParams &tp, PlannedItem &out, std::string &err);
// Tile-shape class key of a planned item (dtype, fused rank, path, size, update).
std::string class_key(const PlannedItem &q);
// Fill a device entry for a planned item (tile_begin left to the caller).
void make_entry(const PlannedItem &q, const BatchItemSpec &it, BatchEntry &e);
// LDS per block a batch launch uses for entries needing lds_bytes (a
// 16-byte multiple; a launch takes the largest of its entries').
uint32_t batch_lds_bytes(uint32_t lds_bytes);
// LDS class of an entry (8 KiB steps): split-mode launches group by it.
uint32_t batch_lds_class(uint32_t lds_bytes);
// Blocks of one launch.  The kernel walks the launch's tiles either by a
// static stride (block b takes slots b, b + grid, ...) or, with a dynamic
// queue, by taking chunks off a device counter.  By default it is launched
// with one block per tile, so the hardware dispatcher hands out the tiles in
// order as blocks retire: measured faster on MI355X than a persistent grid
// walking a static stride, whose blocks drift apart and lose the in-order
// sweep (fp64 manifest of 4096 tensors: 9.2 ms against 10.6 ms).  With
// `persistent` the grid is the resident-block count (occupancy of the
// (dtype, variant) kernel at lds_bytes, times the CUs; looked up once and
// cached).  HPTTX_BATCH_GRID=persistent makes batch_persistent() true.
bool batch_persistent();
uint32_t batch_grid(int dtype, uint32_t variant, uint32_t lds_bytes, uint32_t total_tiles,
                    bool persistent);
// A dynamic work queue over one launch's tiles: the tiles (largest
// transposition first) cut into chunks that shrink towards the end (each a
// quarter of the remaining tiles' share per block, at least one tile), and a
// device counter from which every block takes the index of its next chunk
// with one atomic increment.  A kernel node clears the counter before each
// launch (no memset node, so the sequence can be captured and replayed).
struct BatchQueue {
  uint32_t *counter = nullptr;        // device, one uint32
  const uint32_t *chunks = nullptr;   // device, nchunks + 1 chunk boundaries
  uint32_t nchunks = 0;
};
std::vector<uint32_t> batch_queue_chunks(uint32_t total_tiles, uint32_t grid);
// Launch the batch kernel over entries [0, nent) whose tiles number
// `total`, on `grid` blocks (a 2-D grid beyond 2^20 blocks).  `variant` is
// a variant code (every entry has it) or batch_group_key(., true) for the
// instance that dispatches per entry.
// d_ent_of: null (each block binary-searches d_begins for its tile's entry)
// or the entry of every tile (batch_entry_of_tile: one load per block).
// queue: null (static stride) or a dynamic queue; its counter is cleared
// by a kernel launched just before, on the same stream.
hipError_t launch_batch_group(int dtype, uint32_t variant, const BatchEntry *d_ent,
                              const uint32_t *d_begins, const uint32_t *d_ent_of, uint32_t nent,
                              uint32_t total, uint32_t lds_bytes, uint32_t grid, hipStream_t st,
                              const BatchQueue *queue = nullptr);
// Entry index of each of `total` tiles for entries whose first tiles are
// `begins`.
std::vector<uint32_t> batch_entry_of_tile(const std::vector<uint32_t> &begins, uint32_t total);
// Kernel key of an entry's group: its variant, or the per-type instance
// that dispatches variants at run time when `merged`.
uint32_t batch_group_key(uint32_t variant, bool merged);
// Grouping: one launch per element type (merged) or one per (element type,
// variant) (split), each further split by LDS class.  The environment
// variable HPTTX_BATCH_MODE ("merged" / "split") overrides `fallback`;
// batches and the server default to split (measured faster for both: each
// instance holds only its own variant's registers).
bool batch_merged_default(bool fallback);

// Batch layouts (hpttx_batch_create_ex flags).
//   split (default): one launch per (element type, variant, LDS class), one
//     block per tile, forked over side streams.
//   one launch: one persistent launch per element type whose blocks take
//     chunks of tiles off a dynamic queue (the kernel dispatches each
//     entry's variant at run time).
constexpr uint32_t kBatchOneLaunch = 1;

// The entries of one launch: one element type, and one variant unless merged.
struct BatchGroup {
  int dtype = 0;
  uint32_t variant = 0;             // kernel key (batch_group_key)
  uint32_t nent = 0;
  uint32_t total_tiles = 0;
  uint32_t lds_bytes = 0;
  uint32_t grid = 0;
  BatchEntry *d_ent = nullptr;      // device table
  uint32_t *d_begins = nullptr;     // device: first tile of each entry
  uint32_t *d_ent_of = nullptr;     // device: entry of each tile
  uint32_t *d_queue = nullptr;      // device: dynamic-queue counter (one-launch layout)
  uint32_t *d_chunks = nullptr;     // device: its chunk boundaries
  uint32_t nchunks = 0;             // 0: static walk
  std::vector<BatchEntry> h_ent;    // host copy
  std::vector<uint32_t> order;      // entry k <- user item order[k]
  uint64_t bytes = 0;               // algorithmic bytes
};

struct Batch {
  std::vector<BatchGroup> groups;
  uint32_t count = 0;
  hipStream_t stream = nullptr;
  // Launches fork from the exec stream onto side streams and join back, so
  // the groups run concurrently (also inside the captured graph).
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> events;   // [0] fork, [1 + i] join of side[i]
  std::vector<int> lane;            // stream of each group: 0 exec, i+1 side[i]
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // Recorded after every exec, on whichever stream it ran: set_data and
  // destroy wait on it before they touch the device tables.
  hipEvent_t last_exec = nullptr;
  bool use_graph = true;
  uint64_t bytes = 0;
  std::vector<std::string> paths;   // per user item
  std::vector<uint32_t> item_group; // per user item: its group
  std::vector<uint32_t> item_pos;   // per user item: its entry within the group
};

// Plan every item (tile shapes tuned per class when tuning_timeout != 0)
// and lay the launches out as `flags` asks (kBatchOneLaunch or 0).
int batch_build(const std::vector<BatchItemSpec> &items, hipStream_t stream, double tuning_timeout,
                uint32_t flags, Batch &b, std::string &err);
// Enqueue one kernel launch per group, forked over the side streams.
hipError_t batch_enqueue(Batch &b, hipStream_t stream);
// Capture batch_enqueue into a graph (done on first exec).
hipError_t batch_capture(Batch &b);
void batch_free(Batch &b);
// Re-point item k (user order) at new buffers; uploads the table entry.
int batch_set_data(Batch &b, uint32_t k, const void *A, void *B, std::string &err);

}  // namespace hpttx
