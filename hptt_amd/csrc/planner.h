// planner.h -- host-side problem normalisation and launch planning.
//
// Replaces the reference's problem model and planner
// (inc/hptt/param/parameter_trans.tcc:103-283 dimension merging and leading-dim
// strides; inc/hptt/plan/plan_trans_util.tcc loop-order / thread heuristics)
// with: dimension fusion on strides, a path choice (TILE / ROWS), the
// TILE super-dim groups and tile shape, and vector widths.
#pragma once

#include <stdint.h>

#include <string>

#include "hpttx_core.h"

namespace hpttx {

// Tile-shape knobs the autotuner searches (0 = model default).
struct TuneParams {
  uint32_t run_r = 0;        // target read-run length along A, in elements
  uint32_t run_c = 0;        // target write-run length along B, in elements
  uint32_t block_elems = 0;  // target elements per block tile
  uint32_t wide_min_bytes = 0;  // stride-1-kept rows at least this long use ROWS
  uint32_t stream = 0;       // non-temporal accesses: 0 = model, 1 = loads and stores,
                             // 2 = none, 3 = loads only, 4 = stores only, 5 = stores and
                             // B's loads (rows take 3 and 4 as 1)
  bool operator==(const TuneParams &o) const {
    return run_r == o.run_r && run_c == o.run_c && block_elems == o.block_elems &&
           wide_min_bytes == o.wide_min_bytes && stream == o.stream;
  }
};

struct FusedDim {
  uint64_t n;
  int64_t sa, sb;
};

// The problem after validation, size-1 squeeze and fusion.  Dims are kept in
// A order (ascending sa); border lists them in B order (ascending sb).
struct Problem {
  int dtype = 0;
  uint32_t esize = 4;
  uint32_t rank = 0;          // user rank
  uint32_t in_size[kMaxRank] = {0};
  uint32_t perm[kMaxRank] = {0};
  uint32_t in_outer[kMaxRank] = {0};
  uint32_t out_outer[kMaxRank] = {0};
  int fr = 0;                 // fused rank
  FusedDim d[kMaxRank];
  int border[kMaxRank];
  uint64_t N = 0;             // elements moved
  uint64_t spanA = 0, spanB = 0;  // elements each buffer must hold
};

enum Kind : int { kTile = 0, kRows = 1 };

// MI355X's last-level cache (Infinity Cache): problems whose A and B exceed
// it stream from HBM, and the planner's cache-related choices (streaming
// mode, line-aligned runs, XCD-aware tile order) apply to them only.
constexpr uint64_t kLastLevelCacheBytes = 256ull << 20;

// Tiles whose LDS footprint lets at most 4 blocks share a CU (more than
// 32 KiB of its 160 KiB) run 512-thread blocks, so those 4 blocks still fill
// all 32 wave slots; smaller tiles keep kBlock threads and more blocks.
constexpr uint32_t kWideBlock = 512;
constexpr size_t kWideBlockMinLds = 32768;

struct LaunchPlan {
  Kind kind = kRows;
  TileDesc tile{};
  RowDesc rows{};
  uint32_t vr_shape = 1, vw_shape = 1;  // widest vectors the strides allow
  uint32_t grid = 0;
  size_t lds_bytes = 0;
  uint32_t threads = kBlock;   // threads per block (kBlock or kWideBlock)
  bool nt_load = false;        // non-temporal loads of A
  bool nt_load_b = false;      // non-temporal loads of B (when updating)
  bool nt_store = false;       // non-temporal stores
  TuneParams tune{};
  std::string path;            // "tile", "tile-shared", "wide-copy", "axpby", "gather"
};

// Status codes shared with the C ABI.
enum Status : int {
  kOk = 0,
  kEInval = 1,
  kENoMem = 2,
  kEHip = 3,
  kEUnsupported = 4,
};

uint32_t dtype_esize(int dtype);

// Validation (src/hptt.cc:25-56) + normalisation.  On error, `err` says why.
int build_problem(int dtype, uint32_t rank, const uint32_t *in_size,
                  const uint32_t *perm, const uint32_t *in_outer,
                  const uint32_t *out_outer, Problem &p, std::string &err);

// Blocks a single plan's launch should have at least (4 per CU): small
// problems get shorter tile runs rather than an under-filled GPU.  Batches
// pass 0 (their launches are filled by many transpositions).
constexpr uint32_t kMinPlanTiles = 1024;

// Shape plan (path, tile or row decomposition, exact-vector widths the
// strides allow); tile runs shrink until there are min_tiles tiles or the
// runs reach 128 / 256 bytes.
int plan_launch(const Problem &p, const TuneParams &tp, LaunchPlan &lp,
                std::string &err, uint32_t min_tiles = 0);

// Re-derive the vector-dependent fields of lp: vr / vw equal to the full
// width 16 / esize select exact vectors on that side, anything less
// unaligned full-width vectors (scalars for 16-byte elements).
void finalize_vectors(const Problem &p, LaunchPlan &lp, uint32_t vr, uint32_t vw);

// Streaming (non-temporal) accesses of a tile plan (plan_launch gives the
// measurements): where every run starts and ends on a 128-byte line, every
// access, except A's loads when each buffer fits in the last-level cache (A
// then stays there for the next call); where some run does not (it leaves
// the rest of its line to a neighbouring tile), loads only (4-byte
// elements) or nothing when each buffer fits, and for larger buffers
// everything (4-byte elements whose runs start and end on 64-byte half
// lines) or nothing.
// tile_whole_lines: the shape's part (run lengths and the strides that place
// run starts); stream_for_pointers: the same choice for pointers that are
// not 128-byte aligned (a tuned choice is kept).
bool tile_whole_lines(const Problem &p, const TileDesc &t);
void stream_for_pointers(const Problem &p, LaunchPlan &lp, const void *A, const void *B);

// Widest vector (elements) a pointer's alignment allows for this element size.
uint32_t align_vec(const void *ptr, uint32_t esize, uint32_t cap);

// Blocks per grid row; a launch covers its tiles with ceil(tiles / kGridX)
// rows, one block per tile.
constexpr uint32_t kGridX = 1u << 20;

std::string describe(const Problem &p, const LaunchPlan &lp);

TuneParams default_tune(const Problem &p);

}  // namespace hpttx
