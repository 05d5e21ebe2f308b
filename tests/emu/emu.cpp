// emu.cpp -- TEST INFRASTRUCTURE: host emulator of the library's kernels.
//
// Plans a transposition with the library's own planner (planner.cpp) and runs
// the library's own block bodies (hpttx_core.h) on the host, one emulated
// thread at a time and phase by phase (the kernels separate the same phases
// with __syncthreads()).  Every global and LDS access goes through a bounds
// check, and every store into B is counted, so a test can show that each
// launch shape stays inside its buffers and writes each element of B exactly
// once before any kernel runs on a GPU.  Built with g++ and contraction off.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace emu {
struct Range {
  const char *lo, *hi;
};
static Range g_A, g_B, g_L;
static std::vector<uint8_t> *g_wcount = nullptr;
static uint32_t g_esize = 4;
static std::string g_fault;
static void fault(const char *what, size_t n, bool st) {
  if (g_fault.empty()) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "%s %s of %zu bytes", what, st ? "store" : "load", n);
    g_fault = buf;
  }
}
static void access(const void *p, size_t n, bool st) {
  const char *c = static_cast<const char *>(p);
  auto in = [&](const Range &r) { return c >= r.lo && c + n <= r.hi; };
  if (in(g_L)) {
    // LDS vector accesses must be naturally aligned
    if ((size_t)(c - g_L.lo) % n) fault("misaligned LDS", n, st);
    return;
  }
  // global accesses of a dword or more need only dword alignment
  if (n >= 4 && (reinterpret_cast<uintptr_t>(c) % 4)) fault("non-dword-aligned global", n, st);
  if (!st && in(g_A)) return;
  if (in(g_B)) {
    if (st && g_wcount) {
      size_t first = (size_t)(c - g_B.lo) / g_esize;
      size_t cnt = n / g_esize;
      for (size_t k = 0; k < cnt; ++k) {
        uint8_t &w = (*g_wcount)[first + k];
        if (w < 255) ++w;
      }
    }
    return;
  }
  fault("out-of-bounds", n, st);
}
}  // namespace emu

#define HPTTX_ACCESS(ptr, bytes, is_store) ::emu::access((ptr), (bytes), (is_store))

#include "hpttx_core.h"
#include "planner.h"

using namespace hpttx;

namespace {

template <typename T, typename S, int NC, int VR, int VW, bool UR, bool UW, bool UPD, bool CC,
          bool SH, uint32_t NTH>
void run_tile(const LaunchPlan &lp, const T *A, T *B, const Coef<S> &k) {
  const TileDesc &d = lp.tile;
  std::vector<uint64_t> smem((lp.lds_bytes + 7) / 8 + 2);
  emu::g_L = {reinterpret_cast<const char *>(smem.data()),
              reinterpret_cast<const char *>(smem.data()) + lp.lds_bytes};
  T *lds = reinterpret_cast<T *>(smem.data());
  const TileTabs tb = tile_tabs(d, smem.data(), sizeof(T));
  for (uint32_t t = 0; t < d.ntiles; ++t) {
    const TileCoord tc = tile_coord(d, t);
    // poison LDS (NaN bytes) so that a read of a slot the read pass did not
    // fill shows up in B
    std::memset(smem.data(), 0xff, lp.lds_bytes);
    for (uint32_t tid = 0; tid < NTH; ++tid) tile_tables<NTH>(d, tc, tid, tb);
    // the single-plan read pass, then the batch kernels' (groups of
    // unroll_for loads): both fill the same LDS slots, and both are checked
    for (uint32_t tid = 0; tid < NTH; ++tid)
      tile_read<T, VR, false, kReadGroupPlan, UR, NTH>(d, tc, tid, A, lds, tb);
    for (uint32_t tid = 0; tid < NTH; ++tid)
      tile_read<T, VR, false, unroll_for<T, VR, kUcapBatch>(), UR, NTH>(d, tc, tid, A, lds, tb);
    for (uint32_t tid = 0; tid < NTH; ++tid) {
      if constexpr (quad_write<T, UW, SH>())
        tile_write_quad<T, S, NC, UPD, CC, false, NTH>(d, tc, tid, B, lds, tb, k);
      else
        tile_write<T, S, NC, VW, UPD, CC, SH, false, kUcapPlan, UW, NTH>(d, tc, tid, B, lds, tb,
                                                                        k);
    }
    if (!emu::g_fault.empty()) return;
  }
}

template <typename T, typename S, int NC, int V, bool UA, bool UPD, bool CC>
void run_rows(const LaunchPlan &lp, const T *A, T *B, const Coef<S> &k) {
  const RowDesc &d = lp.rows;
  std::vector<int64_t> tabA(d.RB), tabB(d.RB);
  emu::g_L = {nullptr, nullptr};
  for (uint32_t t = 0; t < d.ntiles; ++t) {
    const RowCoord rc = row_coord(d, t);
    for (uint32_t tid = 0; tid < kBlock; ++tid) row_tables(d, rc, tid, tabA.data(), tabB.data());
    for (uint32_t tid = 0; tid < kBlock; ++tid)
      row_body<T, S, NC, V, UPD, CC, false, kUcapPlan, UA>(d, rc, tid, A, B, tabA.data(),
                                                           tabB.data(), k);
    if (!emu::g_fault.empty()) return;
  }
}

// The same variant choice as the library's launch (kernels.hip launch_typed).
template <typename T, typename S, int NC, int VCAP, bool UPD, bool CC>
void run_typed(const LaunchPlan &lp, const void *A, void *B, const Coef<S> &k) {
  const T *a = static_cast<const T *>(A);
  T *b = static_cast<T *>(B);
  constexpr bool kUA = VCAP > 1;
  if (lp.kind == kTile) {
    const TileDesc &d = lp.tile;
    const bool sh = d.VS > 1, ur = d.ua_r != 0, uw = d.ua_w != 0;
#define EMU_TILE(UR_, UW_, SH_)                                                        \
  do {                                                                                \
    if (lp.threads == kWideBlock)                                                     \
      run_tile<T, S, NC, VCAP, VCAP, UR_, UW_, UPD, CC, SH_, kWideBlock>(lp, a, b, k); \
    else                                                                              \
      run_tile<T, S, NC, VCAP, VCAP, UR_, UW_, UPD, CC, SH_, kBlock>(lp, a, b, k);     \
  } while (0)
    if (sh) {
      if (!ur && !uw) EMU_TILE(false, false, true);
      else if (!ur) EMU_TILE(false, kUA, true);
      else if (!uw) EMU_TILE(kUA, false, true);
      else EMU_TILE(kUA, kUA, true);
    } else {
      if (!ur && !uw) EMU_TILE(false, false, false);
      else if (!ur) EMU_TILE(false, kUA, false);
      else if (!uw) EMU_TILE(kUA, false, false);
      else EMU_TILE(kUA, kUA, false);
    }
#undef EMU_TILE
  } else {
    const RowDesc &d = lp.rows;
    if (d.v == VCAP && VCAP > 1 && !d.ua) run_rows<T, S, NC, VCAP, false, UPD, CC>(lp, a, b, k);
    else if (d.v == VCAP && VCAP > 1) run_rows<T, S, NC, VCAP, kUA, UPD, CC>(lp, a, b, k);
    else run_rows<T, S, NC, 1, false, UPD, CC>(lp, a, b, k);
  }
}

template <typename T, typename S, int NC, int VCAP, bool CCOK>
void run_upd(const LaunchPlan &lp, const void *A, void *B, const double *c, bool upd, bool cc) {
  Coef<S> k{(S)c[0], (S)c[1], (S)c[2], (S)c[3]};
  if (cc && CCOK) {
    if (upd) run_typed<T, S, NC, VCAP, true, CCOK>(lp, A, B, k);
    else run_typed<T, S, NC, VCAP, false, CCOK>(lp, A, B, k);
  } else {
    if (upd) run_typed<T, S, NC, VCAP, true, false>(lp, A, B, k);
    else run_typed<T, S, NC, VCAP, false, false>(lp, A, B, k);
  }
}

void put(char *buf, size_t len, const std::string &s) {
  if (!buf || !len) return;
  size_t n = s.size() < len - 1 ? s.size() : len - 1;
  std::memcpy(buf, s.data(), n);
  buf[n] = 0;
}

}  // namespace

extern "C" {

// Plans like the library (same planner, same pointer-alignment rule) and
// emulates the launch.  tune: NULL or {run_r, run_c, block_elems, wide_min, stream}.
// Returns 0 on success, the planner's status on a planning error, 90 on an
// out-of-bounds access, 91 when an element of B inside the tensor is not
// written exactly once or padding is written.
int emu_transpose(int dtype, const void *A, uint64_t nA, void *B, uint64_t nB, uint32_t rank,
                  const uint32_t *in_size, const uint32_t *perm, double ar, double ai,
                  double br, double bi, int cc, const uint32_t *in_outer,
                  const uint32_t *out_outer, const uint32_t *tune, int force_scalar,
                  char *errbuf, size_t errlen, char *desc, size_t desclen) {
  Problem p;
  std::string err;
  int rc = build_problem(dtype, rank, in_size, perm, in_outer, out_outer, p, err);
  if (rc) { put(errbuf, errlen, err); return rc; }
  TuneParams tp{};
  if (tune) tp = TuneParams{tune[0], tune[1], tune[2], tune[3], tune[4]};
  LaunchPlan lp;
  // explicit tile parameters are taken as given (no small-problem shrink)
  rc = plan_launch(p, tp, lp, err, tune ? 0 : kMinPlanTiles);
  if (rc) { put(errbuf, errlen, err); return rc; }
  // Same rule as the library's exec: exact full vectors where the shape and
  // the pointers' 16-byte alignment allow them, unaligned vectors otherwise
  // (force_scalar: unaligned everywhere, scalars for 16-byte elements).
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
  if (force_scalar) vr = vw = 1;
  finalize_vectors(p, lp, vr, vw);
  put(desc, desclen, describe(p, lp));
  const uint32_t es = p.esize;
  if (nA < p.spanA || nB < p.spanB) { put(errbuf, errlen, "buffers smaller than the tensors"); return 1; }
  emu::g_A = {static_cast<const char *>(A), static_cast<const char *>(A) + nA * es};
  emu::g_B = {static_cast<const char *>(B), static_cast<const char *>(B) + nB * es};
  emu::g_esize = es;
  std::vector<uint8_t> wc(nB, 0);
  emu::g_wcount = &wc;
  emu::g_fault.clear();
  const double c[4] = {ar, ai, br, bi};
  const bool ccx = cc && (ai != 0.0 || bi != 0.0);
  const bool upd = ccx ? (__builtin_fabs(br) >= 1e-16 || __builtin_fabs(bi) >= 1e-16)
                       : (__builtin_fabs(br) >= 1e-16);
  switch (dtype) {
    case kF32: run_upd<Elem<float, 1>, float, 1, 4, false>(lp, A, B, c, upd, false); break;
    case kF64: run_upd<Elem<double, 1>, double, 1, 2, false>(lp, A, B, c, upd, false); break;
    case kC64: run_upd<Elem<float, 2>, float, 2, 2, true>(lp, A, B, c, upd, ccx); break;
    case kC128: run_upd<Elem<double, 2>, double, 2, 1, true>(lp, A, B, c, upd, ccx); break;
  }
  emu::g_wcount = nullptr;
  if (!emu::g_fault.empty()) { put(errbuf, errlen, emu::g_fault); return 90; }
  // Every element of B inside the tensor written exactly once, padding never.
  uint32_t osz[kMaxRank], oo[kMaxRank];
  for (uint32_t d = 0; d < rank; ++d) {
    osz[d] = in_size[perm[d]];
    oo[d] = out_outer ? out_outer[d] : osz[d];
  }
  uint64_t idx[kMaxRank] = {0};
  std::vector<uint8_t> inside(nB, 0);
  for (uint64_t n = 0; n < p.N; ++n) {
    uint64_t off = 0, st = 1;
    for (uint32_t d = 0; d < rank; ++d) { off += idx[d] * st; st *= oo[d]; }
    inside[off] = 1;
    for (uint32_t d = 0; d < rank; ++d) { if (++idx[d] < osz[d]) break; idx[d] = 0; }
  }
  for (uint64_t k = 0; k < nB; ++k) {
    if ((inside[k] && wc[k] != 1) || (!inside[k] && wc[k] != 0)) {
      char buf[160];
      std::snprintf(buf, sizeof buf, "B element %llu written %u times (inside=%d)",
                    (unsigned long long)k, wc[k], inside[k]);
      put(errbuf, errlen, buf);
      return 91;
    }
  }
  return 0;
}

// The planner alone: fills desc and returns the status (no buffers).
int emu_plan(int dtype, uint32_t rank, const uint32_t *in_size, const uint32_t *perm,
             const uint32_t *in_outer, const uint32_t *out_outer, const uint32_t *tune,
             char *errbuf, size_t errlen, char *desc, size_t desclen) {
  Problem p;
  std::string err;
  int rc = build_problem(dtype, rank, in_size, perm, in_outer, out_outer, p, err);
  if (rc) { put(errbuf, errlen, err); return rc; }
  TuneParams tp{};
  if (tune) tp = TuneParams{tune[0], tune[1], tune[2], tune[3], tune[4]};
  LaunchPlan lp;
  rc = plan_launch(p, tp, lp, err, kMinPlanTiles);
  if (rc) { put(errbuf, errlen, err); return rc; }
  put(desc, desclen, describe(p, lp));
  return 0;
}

// Exhaustive-ish check of the fast division against '/' (returns the first
// failing dividend for divisor d, or -1).
int64_t emu_check_fastdiv(uint32_t d, uint64_t n_lo, uint64_t n_hi) {
  FastDiv f = make_fastdiv(d);
  for (uint64_t n = n_lo; n <= n_hi && n <= 0xFFFFFFFFull; ++n)
    if (fdiv((uint32_t)n, f) != (uint32_t)n / d) return (int64_t)n;
  return -1;
}

}  // extern "C"
