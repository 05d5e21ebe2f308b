// batch_kernels.h -- the batch kernel templates (device code).
//
// Included by one source file per element type (batch_kernels_<type>.hip),
// each of which instantiates every variant of its type and fills a table of
// kernel pointers, so the instances compile in parallel.  The host side of
// the batch (planning, grids, launches) is batch.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "batch.h"
#include "hpttx_core.h"

namespace hpttx {

// Variant bits of a batch entry (BatchEntry::variant).
constexpr uint32_t kVarRows = 1, kVarVr = 2, kVarVw = 4, kVarUpd = 8, kVarSh = 16, kVarCc = 32,
                   kVarNt = 64;
constexpr int kVarCodes = 128;

// Fill tab[0 .. kVarCodes] with the kernel instances of one element type:
// tab[code] runs entries of that variant (null where the code does not exist
// for the type), tab[kVarCodes] dispatches each entry's variant at run time.
void batch_kernels_f32(const void **tab);
void batch_kernels_f64(const void **tab);
void batch_kernels_c64(const void **tab);
void batch_kernels_c128(const void **tab);

namespace bk {

template <typename T, typename S, int NC, int VR, int VW, bool UR, bool UW, bool UPD, bool CC,
          bool SH, bool NT, int UCAP>
__device__ __forceinline__ void tile_one(const TileDesc &d, uint32_t lt, uint32_t tid,
                                         const HPTTX_GLOBAL T *A, HPTTX_GLOBAL T *B,
                                         unsigned char *smem, const Coef<S> &co, bool more) {
  T *lds = reinterpret_cast<T *>(smem);
  const TileTabs tb = tile_tabs(d, smem, sizeof(T));
  const TileCoord tc = tile_coord(d, lt);
  tile_tables(d, tc, tid, tb);
  __syncthreads();
  tile_read<T, VR, NT, unroll_for<T, VR, UCAP>(), UR>(d, tc, tid, A, lds, tb);
  __syncthreads();
  if constexpr (quad_write<T, UW, SH>())
    tile_write_quad<T, S, NC, UPD, CC, NT, kBlock>(d, tc, tid, B, lds, tb, co);
  else
    tile_write<T, S, NC, VW, UPD, CC, SH, NT, UCAP, UW>(d, tc, tid, B, lds, tb, co);
  if (more) __syncthreads();   // LDS is reused by the block's next tile
}

template <typename T, typename S, int NC, int V, bool UA, bool UPD, bool CC, bool NT, int UCAP>
__device__ __forceinline__ void rows_one(const RowDesc &d, uint32_t lt, uint32_t tid,
                                         const HPTTX_GLOBAL T *A, HPTTX_GLOBAL T *B,
                                         unsigned char *smem, const Coef<S> &co, bool more) {
  int64_t *tabA = reinterpret_cast<int64_t *>(smem);
  int64_t *tabB = tabA + d.RB;
  const RowCoord rc = row_coord(d, lt);
  row_tables(d, rc, tid, tabA, tabB);
  __syncthreads();
  row_body<T, S, NC, V, UPD, CC, NT, UCAP, UA>(d, rc, tid, A, B, tabA, tabB, co);
  if (more) __syncthreads();
}

// Variant codes that exist for an element type.  Tiles: kVarVr / kVarVw mark
// exact read / write runs; without them the side uses unaligned full-width
// vectors (scalars when VCAP == 1).  Rows: kVarVr = exact vectors, kVarVw =
// unaligned vectors, neither = scalars (strided rows or VCAP == 1); rows
// ignore the shared bit.  Scalar-only types (VCAP == 1) have no vector bits;
// complex coefficients only for complex types.
template <int VCAP, bool CCOK>
constexpr bool valid_code(int c) {
  return !((c & kVarRows) && (c & kVarSh)) &&
         !((c & kVarRows) && (c & kVarVr) && (c & kVarVw)) &&
         !(VCAP == 1 && (c & (kVarVr | kVarVw))) && !(!CCOK && (c & kVarCc));
}

// UCAP: accesses in flight per thread for narrow accesses (kUcapBatch: the
// batch kernels keep 4, which measured faster on MI355X than 8 -- occupancy
// counts for more than depth in flight across many small transpositions).
template <typename T, typename S, int NC, int VCAP, bool CCOK, int CODE, int UCAP>
__device__ __forceinline__ void run_code(const BatchEntry &e, uint32_t lt, uint32_t tid,
                                         unsigned char *smem, bool more) {
  constexpr bool rows = CODE & kVarRows;
  constexpr bool exact_r = CODE & kVarVr;
  constexpr bool exact_w = CODE & kVarVw;
  constexpr bool upd = CODE & kVarUpd;
  constexpr bool sh = CODE & kVarSh;
  constexpr bool cc = CODE & kVarCc;
  constexpr bool nt = CODE & kVarNt;
  const Coef<S> co{(S)e.ar, (S)e.ai, (S)e.br, (S)e.bi};
  // the entry's pointers are device memory: address them as global
  const HPTTX_GLOBAL T *A = (const HPTTX_GLOBAL T *)e.A;
  HPTTX_GLOBAL T *B = (HPTTX_GLOBAL T *)e.B;
  if constexpr (rows) {
    // exact_w on rows: unaligned vectors
    constexpr int v = (exact_r || exact_w) ? VCAP : 1;
    rows_one<T, S, NC, v, exact_w, upd, cc, nt, UCAP>(e.u.rows, lt, tid, A, B, smem, co, more);
  } else {
    constexpr bool ua = VCAP > 1;
    tile_one<T, S, NC, VCAP, VCAP, ua && !exact_r, ua && !exact_w, upd, cc, sh, nt, UCAP>(
        e.u.tile, lt, tid, A, B, smem, co, more);
  }
}

template <typename T, typename S, int NC, int VCAP, bool CCOK, int CODE>
__device__ __forceinline__ void dispatch(uint32_t v, const BatchEntry &e, uint32_t lt,
                                         uint32_t tid, unsigned char *smem, bool more) {
  if constexpr (CODE < kVarCodes) {
    if constexpr (valid_code<VCAP, CCOK>(CODE)) {
      if (v == CODE) {
        run_code<T, S, NC, VCAP, CCOK, CODE, kUcapBatch>(e, lt, tid, smem, more);
        return;
      }
    }
    dispatch<T, S, NC, VCAP, CCOK, CODE + 1>(v, e, lt, tid, smem, more);
  }
}

// Batch kernel: a work-queue walker over the launch's tiles.  CODE >= 0:
// every entry has that variant (one instance per variant); CODE == kMerged:
// each entry's variant is dispatched at run time (one instance per element
// type).  Two walks:
//  * static (counter == null): block b takes slots b, b + G, b + 2G, ...
//    for a grid of G blocks (one block per slot by default, or a persistent
//    grid of resident blocks); each slot is renumbered to a tile so that one
//    XCD runs neighbouring tiles;
//  * dynamic (counter != null, a persistent grid): each block takes the
//    index of its next chunk of consecutive tiles off a device counter (one
//    atomic increment per chunk), until the chunks run out.
constexpr int kMerged = -1;

// Entry lookup cache of one block: tiles [first, end) belong to entry ei.
struct EntryCursor {
  uint32_t ei = 0, first = 0, end = 0;   // empty: forces a lookup
};

__device__ __forceinline__ void find_entry(const BatchEntry *__restrict__ ents,
                                           const uint32_t *__restrict__ begins,
                                           const uint32_t *__restrict__ ent_of, uint32_t nent,
                                           uint32_t t, EntryCursor &c) {
  if (t >= c.first && t < c.end) return;
  if (ent_of) {
    // the entry of every tile, tabulated at batch creation
    c.ei = __builtin_amdgcn_readfirstlane(ent_of[t]);
  } else {
    // the last entry whose first tile is <= t
    uint32_t lo = 0, hi = nent;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (__builtin_amdgcn_readfirstlane(begins[mid]) <= t) lo = mid;
      else hi = mid;
    }
    c.ei = lo;
  }
  c.first = __builtin_amdgcn_readfirstlane(ents[c.ei].tile_begin);
  c.end = c.first + __builtin_amdgcn_readfirstlane(ents[c.ei].ntiles);
}

template <typename T, typename S, int NC, int VCAP, bool CCOK, int CODE>
__device__ __forceinline__ void run_tile(const BatchEntry &e, uint32_t lt, uint32_t tid,
                                         unsigned char *smem, bool more) {
  if constexpr (CODE == kMerged)
    dispatch<T, S, NC, VCAP, CCOK, 0>(__builtin_amdgcn_readfirstlane(e.variant), e, lt, tid, smem,
                                      more);
  else
    run_code<T, S, NC, VCAP, CCOK, CODE, kUcapBatch>(e, lt, tid, smem, more);
}

template <typename T, typename S, int NC, int VCAP, bool CCOK, int CODE>
__global__ __launch_bounds__(kBlock) void k_batch(const BatchEntry *__restrict__ ents,
                                                  const uint32_t *__restrict__ begins,
                                                  const uint32_t *__restrict__ ent_of,
                                                  uint32_t nent, uint32_t total,
                                                  uint32_t *__restrict__ counter,
                                                  const uint32_t *__restrict__ chunks,
                                                  uint32_t nchunks) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t s_chunk;
  const uint32_t tid = threadIdx.x;
  EntryCursor cur;
  if (counter) {
    for (;;) {
      if (tid == 0) s_chunk = atomicAdd(counter, 1u);
      // also the barrier before the LDS is reused by this chunk's first tile
      __syncthreads();
      const uint32_t j = __builtin_amdgcn_readfirstlane(s_chunk);
      __syncthreads();   // every thread has read s_chunk before it is rewritten
      if (j >= nchunks) return;
      const uint32_t lo = __builtin_amdgcn_readfirstlane(chunks[j]);
      const uint32_t hi = __builtin_amdgcn_readfirstlane(chunks[j + 1]);
      for (uint32_t t = lo; t < hi; ++t) {
        find_entry(ents, begins, ent_of, nent, t, cur);
        run_tile<T, S, NC, VCAP, CCOK, CODE>(ents[cur.ei], t - cur.first, tid, smem, t + 1 < hi);
      }
    }
  }
  const uint32_t nblocks = gridDim.x * gridDim.y;
  for (uint32_t i = blockIdx.y * gridDim.x + blockIdx.x; i < total; i += nblocks) {
    // XCD-aware order (for a grid that is a multiple of 8 or covers every
    // slot): neighbouring tiles share cache lines at their edges, and each
    // XCD has its own L2.
    const uint32_t t = xcd_tile(i, total);
    find_entry(ents, begins, ent_of, nent, t, cur);
    run_tile<T, S, NC, VCAP, CCOK, CODE>(ents[cur.ei], t - cur.first, tid, smem,
                                         i + nblocks < total);
  }
}

// Host table of kernel instances: [dtype][variant code] (null where the code
// does not exist for the type).
template <typename T, typename S, int NC, int VCAP, bool CCOK, int CODE>
void fill_kernels(const void **tab) {
  if constexpr (CODE == 0)
    tab[kVarCodes] = reinterpret_cast<const void *>(&k_batch<T, S, NC, VCAP, CCOK, kMerged>);
  if constexpr (CODE < kVarCodes) {
    if constexpr (valid_code<VCAP, CCOK>(CODE))
      tab[CODE] = reinterpret_cast<const void *>(&k_batch<T, S, NC, VCAP, CCOK, CODE>);
    else
      tab[CODE] = nullptr;
    fill_kernels<T, S, NC, VCAP, CCOK, CODE + 1>(tab);
  }
}

}  // namespace bk
}  // namespace hpttx
