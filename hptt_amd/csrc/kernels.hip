// kernels.hip -- gfx950 kernels for one transposition and their dispatch.
//
// TILE  (k_tile): LDS-staged transpose of an R x C x O block tile; coalesced
//       vector reads along A's stride-1 super-dim, coalesced vector writes
//       along B's.  Replaces the reference's macro/micro kernels
//       (src/kernels/macro_kernel_trans.cc:50-142, src/arch/avx2/
//       kernel_trans_avx2.cc:21-467) and its loop nest + OpenMP split
//       (inc/hptt/operations/operation_trans.tcc:92-118,
//       inc/hptt/cgraph/cgraph_trans.tcc:27-48).
// ROWS  (k_rows): rows copied with alpha/beta; the stride-1-kept "wide copy"
//       (reference MacroTransLinear, macro_kernel_trans.cc:148-195,
//       kernel_trans_avx2.cc:480-530), the identity/axpby case and a strided
//       gather for layouts without a unit-stride dim.
// In streaming mode (chosen per plan by the planner or the tuner) loads and
// stores carry the non-temporal hint.  The reference streams its stores only
// when B is not read (macro_kernel_trans.cc:60-62); on gfx950 the hint pays
// for loads and stores alike once every access is part of a long run and the
// problem is larger than the last-level cache.
#include <hip/hip_runtime.h>

#include "hpttx_core.h"
#include "kernels.h"

namespace hpttx {

// One block of NTH threads per tile on a 2-D grid (t = y * gridDim.x + x).
// UR / UW: unaligned runs on the read / write side (full-width vectors at any
// element offset, partial last vectors element by element).  NT / NTB / NTS:
// the non-temporal hint on A's loads / on B's loads (when updating) / on
// stores.
template <typename T, typename S, int NC, int VR, int VW, bool UR, bool UW, bool UPD, bool CC,
          bool SH, bool NT, bool NTB, bool NTS, uint32_t NTH>
__global__ __launch_bounds__(NTH) void k_tile(TileDesc d, const T *__restrict__ A,
                                              T *__restrict__ B, Coef<S> co) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t i = blockIdx.y * gridDim.x + blockIdx.x;
  if (i >= d.ntiles) return;
  const uint32_t t = d.xcd ? xcd_tile(i, d.ntiles) : i;
  T *lds = reinterpret_cast<T *>(smem);
  const TileTabs tb = tile_tabs(d, smem, sizeof(T));
  const uint32_t tid = threadIdx.x;
  const TileCoord tc = tile_coord(d, t);
  tile_tables<NTH>(d, tc, tid, tb);
  __syncthreads();
  tile_read<T, VR, NT, kReadGroupPlan, UR, NTH>(d, tc, tid, (const HPTTX_GLOBAL T *)A, lds, tb);
  __syncthreads();
  if constexpr (quad_write<T, UW, SH>())
    tile_write_quad<T, S, NC, UPD, CC, NTB, NTH, NTS>(d, tc, tid, (HPTTX_GLOBAL T *)B, lds, tb, co);
  else
    tile_write<T, S, NC, VW, UPD, CC, SH, NTB, kUcapPlan, UW, NTH, NTS>(d, tc, tid,
                                                                       (HPTTX_GLOBAL T *)B, lds, tb, co);
}

// NTA / NT: the non-temporal hint on A's loads / on B's loads and the stores.
template <typename T, typename S, int NC, int V, bool UA, bool UPD, bool CC, bool NTA, bool NT>
__global__ __launch_bounds__(kBlock) void k_rows(RowDesc d, const T *__restrict__ A,
                                                 T *__restrict__ B, Coef<S> co) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t t = blockIdx.y * gridDim.x + blockIdx.x;
  if (t >= d.ntiles) return;
  int64_t *tabA = reinterpret_cast<int64_t *>(smem);
  int64_t *tabB = tabA + d.RB;
  const uint32_t tid = threadIdx.x;
  const RowCoord rc = row_coord(d, t);
  row_tables(d, rc, tid, tabA, tabB);
  __syncthreads();
  row_body<T, S, NC, V, UPD, CC, NT, kUcapPlan, UA, NTA>(d, rc, tid, (const HPTTX_GLOBAL T *)A,
                                                        (HPTTX_GLOBAL T *)B, tabA, tabB, co);
}

namespace {

template <typename T, typename S, int NC, int VCAP, bool UPD, bool CC, bool NT, bool NTB,
          bool NTS>
hipError_t launch_typed(const LaunchPlan &lp, const void *A, void *B,
                        const Coef<S> &co, hipStream_t st) {
  const T *a = static_cast<const T *>(A);
  T *b = static_cast<T *>(B);
  if (lp.grid == 0) return hipSuccess;
  // 2-D grid: x * threads must stay below 2^32 work items per dimension.
  const uint32_t gx = lp.grid < kGridX ? lp.grid : kGridX;
  dim3 grid(gx, (lp.grid + gx - 1) / gx), block(kBlock), wide(kWideBlock);
  // Per side: exact vectors, unaligned vectors (VCAP > 1) or scalars (VCAP == 1).
  constexpr bool kUA = VCAP > 1;
  if (lp.kind == kTile) {
    const TileDesc &d = lp.tile;
    const bool sh = d.VS > 1;
    const bool ur = d.ua_r != 0, uw = d.ua_w != 0;
#define HPTTX_TILE(UR_, UW_, SH_)                                                              \
  do {                                                                                        \
    if (lp.threads == kWideBlock)                                                             \
      hipLaunchKernelGGL((k_tile<T, S, NC, VCAP, VCAP, UR_, UW_, UPD, CC, SH_, NT, NTB, NTS, kWideBlock>), \
                         grid, wide, lp.lds_bytes, st, d, a, b, co);                          \
    else                                                                                      \
      hipLaunchKernelGGL((k_tile<T, S, NC, VCAP, VCAP, UR_, UW_, UPD, CC, SH_, NT, NTB, NTS, kBlock>), \
                         grid, block, lp.lds_bytes, st, d, a, b, co);                         \
  } while (0)
    if (sh) {
      if (!ur && !uw) HPTTX_TILE(false, false, true);
      else if (!ur) HPTTX_TILE(false, kUA, true);
      else if (!uw) HPTTX_TILE(kUA, false, true);
      else HPTTX_TILE(kUA, kUA, true);
    } else {
      if (!ur && !uw) HPTTX_TILE(false, false, false);
      else if (!ur) HPTTX_TILE(false, kUA, false);
      else if (!uw) HPTTX_TILE(kUA, false, false);
      else HPTTX_TILE(kUA, kUA, false);
    }
#undef HPTTX_TILE
  } else {
    const RowDesc &d = lp.rows;
    // rows: A's loads take NT, B's loads and the stores NTS (the planner
    // gives rows all, none, or stores and B's loads)
    if (d.v == VCAP && VCAP > 1 && !d.ua)
      hipLaunchKernelGGL((k_rows<T, S, NC, VCAP, false, UPD, CC, NT, NTS>), grid, block,
                         lp.lds_bytes, st, d, a, b, co);
    else if (d.v == VCAP && VCAP > 1)
      hipLaunchKernelGGL((k_rows<T, S, NC, VCAP, kUA, UPD, CC, NT, NTS>), grid, block,
                         lp.lds_bytes, st, d, a, b, co);
    else
      hipLaunchKernelGGL((k_rows<T, S, NC, 1, false, UPD, CC, NT, NTS>), grid, block,
                         lp.lds_bytes, st, d, a, b, co);
  }
  return hipGetLastError();
}

// Without an update B is not read: those launches take NTB = NT, so that no
// kernel instance differs from another only in a hint it never uses.
template <typename T, typename S, int NC, int VCAP, bool CCOK, bool NT, bool NTB, bool NTS>
hipError_t launch_cc(const LaunchPlan &lp, const void *A, void *B, const Coef<S> &co,
                     bool upd, bool cc, hipStream_t st) {
  if (cc && CCOK) {
    if (upd) return launch_typed<T, S, NC, VCAP, true, CCOK, NT, NTB, NTS>(lp, A, B, co, st);
    return launch_typed<T, S, NC, VCAP, false, CCOK, NT, NT, NTS>(lp, A, B, co, st);
  }
  if (upd) return launch_typed<T, S, NC, VCAP, true, false, NT, NTB, NTS>(lp, A, B, co, st);
  return launch_typed<T, S, NC, VCAP, false, false, NT, NT, NTS>(lp, A, B, co, st);
}

template <typename T, typename S, int NC, int VCAP, bool CCOK>
hipError_t launch_upd(const LaunchPlan &lp, const void *A, void *B, const CoefD &c,
                      bool upd, bool cc, hipStream_t st) {
  Coef<S> co{(S)c.ar, (S)c.ai, (S)c.br, (S)c.bi};
  // Non-temporal hints (A's loads, B's loads, stores): all, none, or stores
  // and B's loads; tiles also loads only and stores only.  B's loads carry
  // the hint whenever A's do.
  if (lp.nt_load && lp.nt_store)
    return launch_cc<T, S, NC, VCAP, CCOK, true, true, true>(lp, A, B, co, upd, cc, st);
  if (lp.nt_store && lp.nt_load_b)
    return launch_cc<T, S, NC, VCAP, CCOK, false, true, true>(lp, A, B, co, upd, cc, st);
  if (lp.kind == kTile) {
    if (lp.nt_load)
      return launch_cc<T, S, NC, VCAP, CCOK, true, true, false>(lp, A, B, co, upd, cc, st);
    if (lp.nt_store)
      return launch_cc<T, S, NC, VCAP, CCOK, false, false, true>(lp, A, B, co, upd, cc, st);
  }
  return launch_cc<T, S, NC, VCAP, CCOK, false, false, false>(lp, A, B, co, upd, cc, st);
}

}  // namespace

hipError_t launch_plan(int dtype, const LaunchPlan &lp, const void *A, void *B,
                       const CoefD &c, bool update, bool cc, hipStream_t st) {
  switch (dtype) {
    case kF32: return launch_upd<Elem<float, 1>, float, 1, 4, false>(lp, A, B, c, update, false, st);
    case kF64: return launch_upd<Elem<double, 1>, double, 1, 2, false>(lp, A, B, c, update, false, st);
    case kC64: return launch_upd<Elem<float, 2>, float, 2, 2, true>(lp, A, B, c, update, cc, st);
    case kC128: return launch_upd<Elem<double, 2>, double, 2, 1, true>(lp, A, B, c, update, cc, st);
    default: return hipErrorInvalidValue;
  }
}

uint32_t vec_cap(int dtype) { return 16 / dtype_esize(dtype); }

}  // namespace hpttx
