// hpttx_core.h -- problem descriptors and the per-block bodies of the
// transposition kernels.
//
// Every body here is written once as a host+device function taking an
// explicit thread id, so the HIP kernels (kernels.hip, batch_kernels.h) run
// exactly this code on gfx950 between __syncthreads() barriers, and the
// test-only emulator runs the same code on the host thread by thread to check
// every index against the buffer extents before a kernel ever touches a GPU.
//
// Operation (reference README.md:6-12, inc/hptt/hptt.h:18-22):
//   B[j] = alpha*A[i] (+ beta*B[j]),  j_d = i_{perm[d]}, column-major A and B.
// Arithmetic order (src/arch/avx2/kernel_trans_avx2.cc:67-108,
// inc/hptt/arch/common/common_impl.h:18-33): UPDATE r = fl(fl(b*y) + fl(a*x)),
// otherwise r = fl(a*x) and B is never read.  Contraction is off for every
// translation unit that includes this header.
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define HPTTX_HD __host__ __device__ __forceinline__
#define HPTTX_RESTRICT __restrict__
#else
#define HPTTX_HD inline
#define HPTTX_RESTRICT __restrict__
#endif

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

// A and B are addressed through pointers in the global address space on the
// device, so the backend emits global_load / global_store (a generic pointer
// read from a table gives flat accesses, which also count against the LDS
// wait counter and so serialise LDS waits behind global memory).  Plain
// pointers on the host.
#if defined(__HIP_DEVICE_COMPILE__)
#define HPTTX_GLOBAL __attribute__((address_space(1)))
#else
#define HPTTX_GLOBAL
#endif

// Access hook: the host emulator used by the tests defines it to bounds-check
// and count every global and LDS access; kernels compile it away.
#ifndef HPTTX_ACCESS
#define HPTTX_ACCESS(ptr, bytes, is_store) ((void)0)
#endif

namespace hpttx {

constexpr int kMaxRank = 6;
constexpr uint32_t kBlock = 256;   // threads per block (4 wave64)
// Cap on accesses in flight per thread for narrow (sub-16-byte) accesses:
// single-plan kernels afford 8; the batch kernels keep 4 (higher occupancy
// measured faster over many small transpositions).
constexpr int kUcapPlan = 8;
constexpr int kUcapBatch = 4;
// Loads in flight per thread in the single-plan tile kernel's read pass.
constexpr int kReadGroupPlan = 1;
// Streaming mode (a per-plan choice): loads and stores of A and B carry the
// non-temporal hint.  It pays when every access is part of a long run (a
// plain copy gains ~10% on gfx950); with short runs, lines shared by
// neighbouring tiles are better kept in cache, so plans choose per shape.

enum Dtype : int { kF32 = 0, kF64 = 1, kC64 = 2, kC128 = 3 };

// ---------------------------------------------------------------------------
// Division by a run-time constant: q = (mulhi(n, m) + n) >> s, exact for every
// 32-bit n (round-up multiplier, Granlund & Montgomery 1994, 64-bit add).
struct FastDiv {
  uint32_t d, m, s, pad_;
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f{};
  f.d = d ? d : 1;
  uint32_t s = 0;
  while ((uint64_t(1) << s) < f.d) ++s;
  f.s = s;
  f.m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << s) - f.d)) / f.d + 1);
  return f;
}

HPTTX_HD uint32_t fdiv(uint32_t n, const FastDiv &f) {
  uint64_t hi = ((uint64_t)n * f.m) >> 32;
  return (uint32_t)((hi + (uint64_t)n) >> f.s);
}

// ---------------------------------------------------------------------------
// A mixed-radix index: super-index x over up to kMaxRank dims (dim 0 fastest)
// mapped to element offsets in A (sa) and B (sb).
struct DimList {
  uint32_t nd;
  uint32_t pad_;
  FastDiv div[kMaxRank];   // div[i].d = extent of dim i
  int64_t sa[kMaxRank];
  int64_t sb[kMaxRank];
};

HPTTX_HD void dl_offsets(const DimList &L, uint32_t x, int64_t &oa, int64_t &ob) {
  int64_t a = 0, b = 0;
#pragma unroll
  for (int i = 0; i < kMaxRank; ++i) {
    if (i >= (int)L.nd) break;
    uint32_t dig = x;
    if (i + 1 < (int)L.nd) {
      uint32_t q = fdiv(x, L.div[i]);
      dig = x - q * L.div[i].d;
      x = q;
    }
    a += (int64_t)dig * L.sa[i];
    b += (int64_t)dig * L.sb[i];
  }
  oa = a;
  ob = b;
}

// ---------------------------------------------------------------------------
// TILE path.  The fused problem is split into
//   S: a shared innermost group of VS elements, contiguous in both A and B
//      (VS = 1 unless A's and B's stride-1 dims coincide and are short);
//   R: a super-dim of VR indices contiguous in A right after S (A offset
//      r*VS), scattered in B (offsets from R.sb);
//   C: a super-dim of VC indices contiguous in B right after S (B offset
//      c*VS), scattered in A (offsets from C.sa);
//   O: every other dim, VO indices, offsets in both from O.
// A block tile is TR x TC x TO of (R, C, O) indices (times VS).  It is read
// with coalesced runs of TR*VS elements along A, staged in LDS as
// [o][c][r*VS+s] rows of lds_row elements, and written with coalesced runs of
// TC*VS elements along B.
struct TileDesc {
  uint32_t VS, VR, TR, VC, TC, VO, TO;
  uint32_t nR, nC, nO;
  uint32_t ntiles;
  uint32_t lds_row;      // LDS row stride in elements (multiple of the read vector)
  uint32_t nvr;          // vectors per read run  (ceil(TR*VS / vr))
  uint32_t nvw;          // vectors per write run (ceil(TC*VS / vw))
  uint32_t vr, vw;       // vector widths in elements
  // Mode bits, packed into one word (the descriptor is passed by value as a
  // kernel argument; small plans are launch-bound, and 16 more bytes of it
  // measured 0.3 us more per launch on a loop of 4096 small plans).
  //  ua_r / ua_w: run mode per side: 0 = exact (every run starts 16-byte
  //    aligned and is a whole number of vectors), 1 = unaligned (vectors at
  //    any element offset, a run's last partial vector element by element).
  //  xcd: tile order; 1 = renumber tiles so that each XCD runs 8 consecutive
  //    ones (set for problems beyond the last-level cache: neighbouring
  //    tiles share rows, lines and DRAM pages, which one XCD then fetches
  //    together).
  //  sh_vec: 1 = shared group of whole write vectors (VS % vw == 0, rows
  //    unswizzled and vector-aligned): the write pass reads each output
  //    vector from LDS whole.
  uint32_t ua_r : 1, ua_w : 1, xcd : 1, sh_vec : 1;
  // LDS chunk swizzle (rows of a power-of-two number of 16-byte chunks, no
  // padding): chunk k of LDS row `row` is stored at k ^ ((row >> log2(W)) &
  // swz_mask), W = 16 / sizeof(element) elements per chunk.  The read pass's vector stores stay conflict-free and the
  // write pass's column reads (consecutive lanes on rows VW apart) spread
  // over the banks; swz_mask = 0 is the plain padded layout.  (A rotation
  // for other row lengths, (k + s) mod n, cut the batched fp32 workload's
  // LDS bank conflicts from 4.5 to 1.9 cycles per LDS instruction but not
  // its time, and its per-access select made small plans 5% slower.)
  uint32_t swz_mask;
  FastDiv div_nR, div_nC, div_nvr, div_TC, div_nvw, div_TR, div_VS;
  FastDiv div_ngw;       // groups of 16 / esize R indices per tile
  DimList R, C, O;
};

// ROWS path.  Rows of L elements with element strides ra (A) and rb (B);
// ra = rb = 1 is the wide-copy case (stride-1 dim kept) and the identity /
// axpby case (a single row).  Row index -> offsets through W, rows walked in
// B order.  A block tile is one segment of SEG elements of RB rows.
struct RowDesc {
  uint32_t L, SEG, nseg, nrows, RB, nrb, ntiles, v, nvs;
  uint32_t ua;           // 1: vectors at any element offset, partial last vector
  int64_t ra, rb;
  FastDiv div_nseg, div_nvs;
  DimList W;
};

template <typename S>
struct Coef {
  S ar, ai, br, bi;
};

// ---------------------------------------------------------------------------
// Element types: NC scalar components of type S (complex = interleaved re/im).
template <typename S, int NC>
struct alignas(sizeof(S) * NC) Elem {
  S c[NC];
};

typedef uint32_t u32x2 __attribute__((vector_size(8)));
typedef uint32_t u32x4 __attribute__((vector_size(16)));
// The same vectors at dword alignment: gfx950 global loads and stores of two
// or four dwords need only a dword-aligned address (the backend emits one
// global_load/store_dwordx2/x4 for them); LDS vector accesses stay aligned.
typedef uint32_t u32x2u __attribute__((vector_size(8), aligned(4)));
typedef uint32_t u32x4u __attribute__((vector_size(16), aligned(4)));

template <int BYTES, bool UA> struct RawVec;
template <> struct RawVec<4, false> { typedef uint32_t type; };
template <> struct RawVec<8, false> { typedef u32x2 type; };
template <> struct RawVec<16, false> { typedef u32x4 type; };
template <> struct RawVec<4, true> { typedef uint32_t type; };
template <> struct RawVec<8, true> { typedef u32x2u type; };
template <> struct RawVec<16, true> { typedef u32x4u type; };

// UA: the address is only element-aligned (global memory only).
template <typename T, int V, bool NT = false, bool UA = false>
HPTTX_HD void vload(const T *p, T (&out)[V]) {
  typedef typename RawVec<sizeof(T) * V, UA>::type R;
  HPTTX_ACCESS(p, sizeof(R), false);
#if defined(__HIP_DEVICE_COMPILE__)
  R r = NT ? __builtin_nontemporal_load(reinterpret_cast<const R *>(p))
           : *reinterpret_cast<const R *>(p);
#else
  R r = *reinterpret_cast<const R *>(p);
#endif
  __builtin_memcpy(&out[0], &r, sizeof(R));
}

template <typename T, int V, bool NT, bool UA = false>
HPTTX_HD void vstore(T *p, const T (&in)[V]) {
  typedef typename RawVec<sizeof(T) * V, UA>::type R;
  R r;
  HPTTX_ACCESS(p, sizeof(R), true);
  __builtin_memcpy(&r, &in[0], sizeof(R));
#if defined(__HIP_DEVICE_COMPILE__)
  if (NT)
    __builtin_nontemporal_store(r, reinterpret_cast<R *>(p));
  else
    *reinterpret_cast<R *>(p) = r;
#else
  *reinterpret_cast<R *>(p) = r;
#endif
}

// The same for A and B in global memory (UA: only element-aligned).
template <typename T, int V, bool NT = false, bool UA = false>
HPTTX_HD void gload(const HPTTX_GLOBAL T *p, T (&out)[V]) {
  typedef typename RawVec<sizeof(T) * V, UA>::type R;
  HPTTX_ACCESS(p, sizeof(R), false);
#if defined(__HIP_DEVICE_COMPILE__)
  R r = NT ? __builtin_nontemporal_load(reinterpret_cast<const HPTTX_GLOBAL R *>(p))
           : *reinterpret_cast<const HPTTX_GLOBAL R *>(p);
#else
  R r = *reinterpret_cast<const R *>(p);
#endif
  __builtin_memcpy(&out[0], &r, sizeof(R));
}

template <typename T, int V, bool NT, bool UA = false>
HPTTX_HD void gstore(HPTTX_GLOBAL T *p, const T (&in)[V]) {
  typedef typename RawVec<sizeof(T) * V, UA>::type R;
  R r;
  HPTTX_ACCESS(p, sizeof(R), true);
  __builtin_memcpy(&r, &in[0], sizeof(R));
#if defined(__HIP_DEVICE_COMPILE__)
  if (NT)
    __builtin_nontemporal_store(r, reinterpret_cast<HPTTX_GLOBAL R *>(p));
  else
    *reinterpret_cast<HPTTX_GLOBAL R *>(p) = r;
#else
  *reinterpret_cast<R *>(p) = r;
#endif
}

// r = fl(a*x) or fl(fl(b*y) + fl(a*x)) per component; complex coefficients
// use ax = (fl(fl(ar*xr) - fl(ai*xi)), fl(fl(ar*xi) + fl(ai*xr))), same for by.
template <typename S, int NC, bool UPD, bool CC>
HPTTX_HD Elem<S, NC> axpby(const Elem<S, NC> &x, const Elem<S, NC> &y,
                           const Coef<S> &k) {
  Elem<S, NC> r;
  if (!CC) {
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      S ax = k.ar * x.c[i];
      if (UPD) {
        S by = k.br * y.c[i];
        r.c[i] = by + ax;
      } else {
        r.c[i] = ax;
      }
    }
  } else {
    S p0 = k.ar * x.c[0];
    S p1 = k.ai * x.c[NC - 1];
    S p2 = k.ar * x.c[NC - 1];
    S p3 = k.ai * x.c[0];
    S axr = p0 - p1;
    S axi = p2 + p3;
    if (UPD) {
      S q0 = k.br * y.c[0];
      S q1 = k.bi * y.c[NC - 1];
      S q2 = k.br * y.c[NC - 1];
      S q3 = k.bi * y.c[0];
      S byr = q0 - q1;
      S byi = q2 + q3;
      r.c[0] = byr + axr;
      r.c[NC - 1] = byi + axi;
    } else {
      r.c[0] = axr;
      r.c[NC - 1] = axi;
    }
  }
  return r;
}

// ---------------------------------------------------------------------------
// TILE body.
struct TileCoord {
  uint32_t r0, c0, o0, rlen, clen, olen;
};

HPTTX_HD uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }

HPTTX_HD TileCoord tile_coord(const TileDesc &d, uint32_t t) {
  TileCoord tc;
  uint32_t q = fdiv(t, d.div_nR);
  uint32_t tr = t - q * d.nR;
  uint32_t to = fdiv(q, d.div_nC);
  uint32_t tcc = q - to * d.nC;
  tc.r0 = tr * d.TR;
  tc.c0 = tcc * d.TC;
  tc.o0 = to * d.TO;
  tc.rlen = umin32(d.TR, d.VR - tc.r0);
  tc.clen = umin32(d.TC, d.VC - tc.c0);
  tc.olen = umin32(d.TO, d.VO - tc.o0);
  return tc;
}

// Position of 16-byte chunk k of LDS row `row` (see TileDesc::swz_mask).
template <typename T>
HPTTX_HD uint32_t lds_chunk(const TileDesc &d, uint32_t row, uint32_t k) {
  constexpr uint32_t kShift = sizeof(T) >= 16 ? 0 : sizeof(T) == 8 ? 1 : 2;
  return k ^ ((row >> kShift) & d.swz_mask);
}

// XCD-aware tile order: blocks are dispatched round-robin over the 8 XCDs,
// so slots i with i % 8 == x run on XCD x.  Renumbering each group of 64
// slots makes those 8 consecutive tiles (a bijection on [0, n); the last
// partial group keeps its order).
HPTTX_HD uint32_t xcd_tile(uint32_t i, uint32_t n) {
  return i < (n & ~63u) ? (i & ~63u) | ((i & 7u) << 3) | ((i >> 3) & 7u) : i;
}

// LDS footprint of a tile, in bytes: data rows then the four offset tables.
inline size_t tile_lds_bytes(const TileDesc &d, size_t esize) {
  size_t data = (size_t)d.TO * d.TC * d.lds_row * esize;
  data = (data + 15) & ~(size_t)15;
  return data + sizeof(int64_t) * ((size_t)d.TC + d.TR + 2 * (size_t)d.TO);
}

struct TileTabs {
  int64_t *AC;  // A offset of C index c0 + c      [TC]
  int64_t *BR;  // B offset of R index r0 + r      [TR]
  int64_t *AO;  // A offset of O index o0 + o      [TO]
  int64_t *BO;  // B offset of O index o0 + o      [TO]
};

HPTTX_HD TileTabs tile_tabs(const TileDesc &d, void *lds_base, size_t esize) {
  size_t data = (size_t)d.TO * d.TC * d.lds_row * esize;
  data = (data + 15) & ~(size_t)15;
  int64_t *p = reinterpret_cast<int64_t *>(reinterpret_cast<char *>(lds_base) + data);
  TileTabs t;
  t.AC = p;
  t.BR = p + d.TC;
  t.AO = p + d.TC + d.TR;
  t.BO = p + d.TC + d.TR + d.TO;
  return t;
}

// NTH: threads per block (every body below takes it; kBlock unless a
// single-plan launch uses wider blocks).
template <uint32_t NTH = kBlock>
HPTTX_HD void tile_tables(const TileDesc &d, const TileCoord &tc, uint32_t tid,
                          const TileTabs &tb) {
  for (uint32_t i = tid; i < tc.clen; i += NTH) {
    int64_t a, b;
    dl_offsets(d.C, tc.c0 + i, a, b);
    tb.AC[i] = a;
  }
  for (uint32_t i = tid; i < tc.rlen; i += NTH) {
    int64_t a, b;
    dl_offsets(d.R, tc.r0 + i, a, b);
    tb.BR[i] = b;
  }
  for (uint32_t i = tid; i < tc.olen; i += NTH) {
    int64_t a, b;
    dl_offsets(d.O, tc.o0 + i, a, b);
    tb.AO[i] = a;
    tb.BO[i] = b;
  }
}

// Independent accesses a thread keeps in flight per step: 4 x 16-byte
// vectors, and up to 8 narrower accesses (32 bytes of scalars), so that
// shapes that only allow narrow accesses keep more bytes in flight without
// giving up occupancy.
template <typename T, int V, int CAP>
constexpr int unroll_for() {
  return (64 / (int)(V * sizeof(T))) < 4 ? 4
       : (64 / (int)(V * sizeof(T))) > CAP ? CAP : (64 / (int)(V * sizeof(T)));
}

// Store the first n <= V elements of a partial vector one by one.
template <typename T, int V, bool NT>
HPTTX_HD void store_part(HPTTX_GLOBAL T *p, uint32_t n, const T (&in)[V]) {
#pragma unroll
  for (int j = 0; j < V; ++j) {
    if ((uint32_t)j < n) {
      const T one[1] = {in[j]};
      gstore<T, 1, NT>(p + j, one);
    }
  }
}

// x[j] = x[j + s] for j + s < V (elements past V - s keep their values, which
// the callers never use).  A partial last vector of n elements loaded as the
// full vector ending at its run's end arrives shifted up by s = V - n.
// Selection is done on 32-bit words (a select between element structs made
// the compiler keep the vector in scratch memory).
template <typename T, int V>
HPTTX_HD void shift_down(T (&x)[V], uint32_t s) {
  constexpr int kW = (int)(sizeof(T) / 4);   // words per element
  uint32_t w[V * kW];
  __builtin_memcpy(&w[0], &x[0], sizeof(w));
#pragma unroll
  for (int j = 0; j < V; ++j) {
#pragma unroll
    for (int t = 1; j + t < V; ++t) {
#pragma unroll
      for (int i = 0; i < kW; ++i) w[j * kW + i] = (s == (uint32_t)t) ? w[(j + t) * kW + i] : w[j * kW + i];
    }
  }
  __builtin_memcpy(&x[0], &w[0], sizeof(w));
}

// One item's V elements at element offset `off` of p, of which the first n
// lie inside the run (n = 0: a masked item, nothing is used).  Loads are
// issued without a branch, so that a thread's loads for all its items are in
// flight together (under a branch per item the compiler waited for the
// previous load before the next: one load in flight per thread).  `first` is
// a vector (WHOLE) or an element (!WHOLE) inside the buffer that masked items
// load instead.  WHOLE: one vector; a partial last vector of an unaligned run
// (UA, n < V) is loaded as the full vector ending at the run's end, inside
// the run whenever the run holds a whole vector, and arrives shifted up by
// V - n (shift_down undoes it).  !WHOLE (runs shorter than one vector):
// element by element, elements past n loading `first`.
template <typename T, int V, bool NT, bool UA, bool WHOLE>
HPTTX_HD void load_item(const HPTTX_GLOBAL T *p, int64_t off, uint32_t n, int64_t first,
                        T (&out)[V]) {
  if (WHOLE) {
    const int64_t at = n ? off - (int64_t)(V - n) : first;
    gload<T, V, NT, UA>(p + at, out);
  } else {
#pragma unroll
    for (int j = 0; j < V; ++j) {
      T one[1];
      gload<T, 1, NT>(p + ((uint32_t)j < n ? off + j : first), one);
      out[j] = one[0];
    }
  }
}

// A thread's items k0, k0 + NTH, ... < items are taken in groups of G (G
// loads in flight together): groups of U while every thread of the wave has
// all U, then at most one group of each smaller power of two, then single
// items.  Items a group does not hold are masked (loads from `first`, no
// stores), so only the wave that holds the pass's end masks any, and a small
// pass does not issue U loads for fewer items.  k0 | 63 is the largest k0 in
// the thread's wave (waves are 64 consecutive threads, and k0 advances by
// multiples of 64).  f(IntC<G>{}, k0) runs one group.
template <int N>
struct IntC {
  static constexpr int value = N;
};

template <int G, uint32_t NTH, typename F>
HPTTX_HD void tail_groups(uint32_t k0, uint32_t items, const F &f) {
  if constexpr (G > 1) {
    if ((k0 | 63u) + (uint32_t)(G - 1) * NTH < items) {
      f(IntC<G>{}, k0);
      k0 += (uint32_t)G * NTH;
    }
    tail_groups<G / 2, NTH>(k0, items, f);
  } else {
    for (; k0 < items; k0 += NTH) f(IntC<1>{}, k0);
  }
}

template <int U, uint32_t NTH, typename F>
HPTTX_HD void for_groups(uint32_t tid, uint32_t items, const F &f) {
  uint32_t k0 = tid;
  if constexpr (U > 1) {
    for (; (k0 | 63u) + (uint32_t)(U - 1) * NTH < items; k0 += (uint32_t)U * NTH)
      f(IntC<U>{}, k0);
    tail_groups<U / 2, NTH>(k0, items, f);
  } else {
    for (; k0 < items; k0 += NTH) f(IntC<1>{}, k0);
  }
}

// Read pass: item k = (row, v), row = o*TC + c; one vector of VR elements
// from A's run r0*VS + [0, rlen*VS) into LDS row `row`.  UA: the run may
// start at any element offset and end in a partial vector; the LDS side stays
// vector-aligned.  Masked items load the tile's first vector.
template <typename T, int VR, bool NT, int U, bool UA, uint32_t NTH, bool WHOLE>
HPTTX_HD void tile_read_pass(const TileDesc &d, const TileCoord &tc, uint32_t tid,
                             const HPTTX_GLOBAL T *HPTTX_RESTRICT A, T *lds,
                             const TileTabs &tb) {
  const uint32_t nvr = d.nvr;
  const uint32_t items = nvr * d.TC * tc.olen;
  const uint32_t run = tc.rlen * d.VS;
  const int64_t abase = (int64_t)tc.r0 * d.VS;
  const int64_t first = abase + tb.AC[0] + tb.AO[0];
  auto group = [&](auto gc, uint32_t k0) {
    constexpr int G = decltype(gc)::value;
    T buf[G][VR];
    uint32_t lidx[G];
    uint32_t n[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const uint32_t k = k0 + (uint32_t)u * NTH;
      const uint32_t row = fdiv(k, d.div_nvr);
      const uint32_t v = k - row * nvr;
      const uint32_t o = fdiv(row, d.div_TC);
      const uint32_t c = row - o * d.TC;
      const bool ok = (k < items) && (c < tc.clen) && (v * VR < run);
      n[u] = !ok ? 0u : UA ? umin32(VR, run - v * VR) : (uint32_t)VR;
      lidx[u] = row * d.lds_row + lds_chunk<T>(d, row, v) * VR;
      const int64_t off = abase + v * VR + tb.AC[ok ? c : 0] + tb.AO[ok ? o : 0];
      load_item<T, VR, NT, UA, WHOLE>(A, off, n[u], first, buf[u]);
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (!n[u]) continue;
      if (UA && WHOLE) shift_down<T, VR>(buf[u], VR - n[u]);
      vstore<T, VR, false>(lds + lidx[u], buf[u]);
    }
  };
  for_groups<U, NTH>(tid, items, group);
}

// U: loads a thread keeps in flight (the group size).  Single-plan tiles use
// 1 (measured on MI355X against groups of 4: 6-D fp32 6010 against 5930
// GB/s, 6-D fp64 6336 against 6267, the same on the c128 tile); the batch
// kernels use groups of 4 (batched fp32 manifest 4.70 against 4.79 ms).
template <typename T, int VR, bool NT, int U, bool UA = false, uint32_t NTH = kBlock>
HPTTX_HD void tile_read(const TileDesc &d, const TileCoord &tc, uint32_t tid,
                        const HPTTX_GLOBAL T *HPTTX_RESTRICT A, T *lds,
                        const TileTabs &tb) {
  if (!UA || tc.rlen * d.VS >= (uint32_t)VR)
    tile_read_pass<T, VR, NT, U, UA, NTH, true>(d, tc, tid, A, lds, tb);
  else
    tile_read_pass<T, VR, NT, U, UA, NTH, false>(d, tc, tid, A, lds, tb);
}

// Write pass: item k = (row, v), row = o*TR + r; VW elements of B's run
// c0*VS + [0, clen*VS) gathered from LDS column r, combined with alpha/beta,
// written as one vector.  A thread's U items are addressed and (when B is
// read) B's U vectors loaded first (load_item: masked items load the tile's
// first write vector; a partial last vector is loaded as the full vector
// ending at its run's end, whose elements before ours, which another thread
// may be writing at the same time, are discarded), then gathered, combined
// and stored.  Masked items gather from LDS row 0 and store nothing.  (B is
// not prefetched before the read pass: holding its values across the read
// pass costs occupancy, and measured slower on MI355X.)  UA: runs start at
// any element offset and may end in a partial vector, which is written
// element by element.
struct WriteAddr {
  int64_t boff;     // B element offset of the vector
  uint32_t lrow0;   // LDS row of (o, c = 0)
  uint32_t x0;      // element of (r, s = 0) within an LDS row
  uint32_t q0;      // first element of the vector within the run
  uint32_t n;       // elements of the vector inside the run (0: masked item)
};

template <bool UA>
HPTTX_HD WriteAddr write_addr(const TileDesc &d, const TileCoord &tc, uint32_t kk,
                              uint32_t vw, const TileTabs &tb) {
  const uint32_t nvw = d.nvw;
  const uint32_t items = nvw * d.TR * tc.olen;
  const uint32_t row = fdiv(kk, d.div_nvw);
  const uint32_t v = kk - row * nvw;
  const uint32_t o = fdiv(row, d.div_TR);
  const uint32_t r = row - o * d.TR;
  const uint32_t run = tc.clen * d.VS;
  const bool ok = (kk < items) && (r < tc.rlen) && (v * vw < run);
  WriteAddr w;
  w.q0 = ok ? v * vw : 0;
  w.n = !ok ? 0u : UA ? umin32(vw, run - w.q0) : vw;
  w.lrow0 = ok ? o * d.TC : 0;
  w.x0 = ok ? r * d.VS : 0;
  w.boff = (int64_t)tc.c0 * d.VS + w.q0 + tb.BR[ok ? r : 0] + tb.BO[ok ? o : 0];
  return w;
}

template <typename T, typename S, int NC, int VW, bool UPD, bool CC, bool SH, bool NT, int UCAP,
          bool UA, uint32_t NTH, bool WHOLE, bool NTS>
HPTTX_HD void tile_write_pass(const TileDesc &d, const TileCoord &tc, uint32_t tid,
                              HPTTX_GLOBAL T *HPTTX_RESTRICT B, const T *lds,
                              const TileTabs &tb, const Coef<S> &k) {
  constexpr int U = unroll_for<T, VW, UCAP>();
  const uint32_t items = d.nvw * d.TR * tc.olen;
  const int64_t bfirst = (int64_t)tc.c0 * d.VS + tb.BR[0] + tb.BO[0];
  auto group = [&](auto gc, uint32_t k0) {
    constexpr int G = decltype(gc)::value;
    WriteAddr a[G];
    T y[G][VW];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      a[u] = write_addr<UA>(d, tc, k0 + (uint32_t)u * NTH, VW, tb);
      if (UPD) load_item<T, VW, NT, UA, WHOLE>(B, a[u].boff, a[u].n, bfirst, y[u]);
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
      T x[VW];
      if (SH && d.sh_vec) {
        // A shared group of whole vectors (VS % VW == 0, unswizzled rows):
        // the vector's VW elements are consecutive s of one LDS row.
        const uint32_t c = fdiv(a[u].q0, d.div_VS);
        const uint32_t s = a[u].q0 - c * d.VS;
        vload<T, VW>(lds + (a[u].lrow0 + c) * d.lds_row + a[u].x0 + s, x);
      } else {
#pragma unroll
        for (int j = 0; j < VW; ++j) {
          // elements past a partial vector's end, and masked items, read
          // element 0 of the run (in the tile, never stored)
          const uint32_t q = (uint32_t)j < a[u].n ? a[u].q0 + j : 0;
          uint32_t c = q, s = 0;
          if (SH) {
            c = fdiv(q, d.div_VS);
            s = q - c * d.VS;
          }
          // LDS element (row lrow0 + c, x0 + s), through the chunk swizzle
          // (chunks of 16 / sizeof(T) elements, the read pass's vectors)
          constexpr uint32_t kChunk = sizeof(T) >= 16 ? 1 : 16 / sizeof(T);
          const uint32_t lrow = a[u].lrow0 + c, xc = a[u].x0 + s;
          const uint32_t xs = lds_chunk<T>(d, lrow, xc / kChunk) * kChunk + xc % kChunk;
          const T *src = lds + lrow * d.lds_row + xs;
          HPTTX_ACCESS(src, sizeof(T), false);
          x[j] = *src;
        }
      }
      if (UPD && UA && WHOLE && a[u].n) shift_down<T, VW>(y[u], VW - a[u].n);
      T out[VW];
#pragma unroll
      for (int j = 0; j < VW; ++j) {
        const Elem<S, NC> &xe = reinterpret_cast<const Elem<S, NC> &>(x[j]);
        const Elem<S, NC> &ye = reinterpret_cast<const Elem<S, NC> &>(y[u][j]);
        Elem<S, NC> re = axpby<S, NC, UPD, CC>(xe, ye, k);
        __builtin_memcpy(&out[j], &re, sizeof(T));
      }
      if (a[u].n == (uint32_t)VW)
        gstore<T, VW, NTS, UA>(B + a[u].boff, out);
      else if (UA && a[u].n)
        store_part<T, VW, NTS>(B + a[u].boff, a[u].n, out);
    }
  };
  for_groups<U, NTH>(tid, items, group);
}

// The whole write pass.  NT: B's loads carry the non-temporal hint; NTS: its
// stores do.
template <typename T, typename S, int NC, int VW, bool UPD, bool CC, bool SH, bool NT, int UCAP,
          bool UA = false, uint32_t NTH = kBlock, bool NTS = NT>
HPTTX_HD void tile_write(const TileDesc &d, const TileCoord &tc, uint32_t tid,
                         HPTTX_GLOBAL T *HPTTX_RESTRICT B, const T *lds, const TileTabs &tb,
                         const Coef<S> &k) {
  if (!UA || tc.clen * d.VS >= (uint32_t)VW)
    tile_write_pass<T, S, NC, VW, UPD, CC, SH, NT, UCAP, UA, NTH, true, NTS>(d, tc, tid, B, lds,
                                                                              tb, k);
  else
    tile_write_pass<T, S, NC, VW, UPD, CC, SH, NT, UCAP, UA, NTH, false, NTS>(d, tc, tid, B, lds,
                                                                               tb, k);
}

// Write pass by W x W blocks (W = 16 / sizeof(T): 4 for 4-byte and 2 for
// 8-byte elements), for exact write runs and VS = 1.  Item (o, g, q): R
// indices r = W*g .. W*g+W-1 (B rows) and C indices c = W*q .. W*q+W-1 (the
// q-th 16-byte vector of each B run).  The thread reads the 16-byte LDS chunk
// g of the W LDS rows c (W ds_read_b128), so the block arrives as x[c][r];
// B's vector for row r is column r of it.  It issues B's W loads first
// (UPDATE), then the LDS reads, then W stores.  Against the general write
// pass this is 1/W of the LDS instructions and of the per-element address
// arithmetic.  Exact writes make every C run a whole number of vectors
// (clen % W == 0); R indices past rlen (a partial last group) are skipped.
template <typename T, typename S, int NC, bool UPD, bool CC, bool NT, uint32_t NTH, bool NTS = NT>
HPTTX_HD void tile_write_quad(const TileDesc &d, const TileCoord &tc, uint32_t tid,
                              HPTTX_GLOBAL T *HPTTX_RESTRICT B, const T *lds,
                              const TileTabs &tb, const Coef<S> &k) {
  constexpr uint32_t W = 16 / sizeof(T);
  static_assert(W == 2 || W == 4, "blocks of 4- or 8-byte elements");
  const uint32_t nq = d.nvw, ng = d.div_ngw.d;
  const uint32_t items = nq * ng * tc.olen;
  for (uint32_t it = tid; it < items; it += NTH) {
    const uint32_t rest = fdiv(it, d.div_nvw);
    const uint32_t q = it - rest * nq;
    const uint32_t o = fdiv(rest, d.div_ngw);
    const uint32_t g = rest - o * ng;
    const uint32_t c0 = W * q, r0 = W * g;
    if (c0 >= tc.clen || r0 >= tc.rlen) continue;
    const uint32_t nr = umin32(W, tc.rlen - r0);
    const int64_t bbase = (int64_t)tc.c0 + c0 + tb.BO[o];
    int64_t boff[W];
    T y[W][W];
#pragma unroll
    for (uint32_t j = 0; j < W; ++j) {
      boff[j] = j < nr ? bbase + tb.BR[r0 + j] : 0;
      if (UPD && j < nr) gload<T, W, NT>(B + boff[j], y[j]);
    }
    T x[W][W];
#pragma unroll
    for (uint32_t m = 0; m < W; ++m) {
      const uint32_t row = o * d.TC + c0 + m;
      const uint32_t chunk = lds_chunk<T>(d, row, g);
      vload<T, W>(lds + row * d.lds_row + chunk * W, x[m]);
    }
#pragma unroll
    for (uint32_t j = 0; j < W; ++j) {
      if (j >= nr) continue;
      T out[W];
#pragma unroll
      for (uint32_t m = 0; m < W; ++m) {
        const Elem<S, NC> &xe = reinterpret_cast<const Elem<S, NC> &>(x[m][j]);
        const Elem<S, NC> &ye = reinterpret_cast<const Elem<S, NC> &>(y[j][m]);
        Elem<S, NC> re = axpby<S, NC, UPD, CC>(xe, ye, k);
        __builtin_memcpy(&out[m], &re, sizeof(T));
      }
      gstore<T, W, NTS>(B + boff[j], out);
    }
  }
}

// The W x W write pass applies: 4- or 8-byte elements, exact writes, VS == 1.
template <typename T, bool UW, bool SH>
HPTTX_HD constexpr bool quad_write() {
  return (sizeof(T) == 4 || sizeof(T) == 8) && !UW && !SH;
}

// ---------------------------------------------------------------------------
// ROWS body.
struct RowCoord {
  uint32_t row0, nr, seg0, slen;
};

HPTTX_HD RowCoord row_coord(const RowDesc &d, uint32_t t) {
  RowCoord rc;
  uint32_t rbk = fdiv(t, d.div_nseg);
  uint32_t seg = t - rbk * d.nseg;
  rc.row0 = rbk * d.RB;
  rc.nr = umin32(d.RB, d.nrows - rc.row0);
  rc.seg0 = seg * d.SEG;
  rc.slen = umin32(d.SEG, d.L - rc.seg0);
  return rc;
}

inline size_t row_lds_bytes(const RowDesc &d) { return sizeof(int64_t) * 2 * (size_t)d.RB; }

HPTTX_HD void row_tables(const RowDesc &d, const RowCoord &rc, uint32_t tid,
                         int64_t *tabA, int64_t *tabB) {
  for (uint32_t i = tid; i < rc.nr; i += kBlock) {
    int64_t a, b;
    dl_offsets(d.W, rc.row0 + i, a, b);
    tabA[i] = a;
    tabB[i] = b;
  }
}

// UA: rows start at any element offset in A and B and the segment may end
// in a partial vector, which is stored element by element.  Loads go through
// load_item (no branch per item: items past the block load the segment's
// first vector; a partial last vector is loaded as the full vector ending at
// the segment's end).  UA rows are contiguous (ra = rb = 1).  NTA: the
// non-temporal hint on A's loads; NT: on B's loads and the stores.
template <typename T, typename S, int NC, int V, bool UPD, bool CC, bool NT, int UCAP, bool UA,
          bool WHOLE, bool NTA>
HPTTX_HD void row_pass(const RowDesc &d, const RowCoord &rc, uint32_t tid,
                       const HPTTX_GLOBAL T *HPTTX_RESTRICT A, HPTTX_GLOBAL T *HPTTX_RESTRICT B,
                       const int64_t *tabA, const int64_t *tabB, const Coef<S> &k) {
  constexpr int U = unroll_for<T, V, UCAP>();
  const uint32_t nvs = d.nvs;
  const uint32_t items = nvs * rc.nr;
  const int64_t afirst = tabA[0] + (int64_t)rc.seg0 * d.ra;
  const int64_t bfirst = tabB[0] + (int64_t)rc.seg0 * d.rb;
  auto group = [&](auto gc, uint32_t k0) {
    constexpr int G = decltype(gc)::value;
    T x[G][V];
    T y[G][V];
    uint32_t n[G];
    int64_t ob[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const uint32_t kk = k0 + (uint32_t)u * kBlock;
      const uint32_t rr = fdiv(kk, d.div_nvs);
      const uint32_t v = kk - rr * nvs;
      const bool ok = (kk < items) && (v * V < rc.slen);
      n[u] = !ok ? 0u : UA ? umin32(V, rc.slen - v * V) : (uint32_t)V;
      const int64_t e = (int64_t)rc.seg0 + (int64_t)(ok ? v : 0) * V;
      const uint32_t rs = ok ? rr : 0;
      ob[u] = tabB[rs] + e * d.rb;
      load_item<T, V, NTA, UA, WHOLE>(A, tabA[rs] + e * d.ra, n[u], afirst, x[u]);
      if (UPD) load_item<T, V, NT, UA, WHOLE>(B, ob[u], n[u], bfirst, y[u]);
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (!n[u]) continue;
      if (UA && WHOLE) {
        shift_down<T, V>(x[u], V - n[u]);
        if (UPD) shift_down<T, V>(y[u], V - n[u]);
      }
      T out[V];
#pragma unroll
      for (int j = 0; j < V; ++j) {
        const Elem<S, NC> &xe = reinterpret_cast<const Elem<S, NC> &>(x[u][j]);
        const Elem<S, NC> &ye = reinterpret_cast<const Elem<S, NC> &>(y[u][j]);
        Elem<S, NC> re = axpby<S, NC, UPD, CC>(xe, ye, k);
        __builtin_memcpy(&out[j], &re, sizeof(T));
      }
      HPTTX_GLOBAL T *pb = B + ob[u];
      if (!UA || n[u] == V)
        gstore<T, V, NT, UA>(pb, out);
      else
        store_part<T, V, NT>(pb, n[u], out);
    }
  };
  // Every group holds U items here, masked ones included: on the row
  // kernel's 2-vector segments this measured 5% faster (6.57 against 6.27
  // TB/s on the 1.8 GB axpby) than groups sized to the items left.
  for (uint32_t k0 = tid; k0 < items; k0 += kBlock * U) group(IntC<U>{}, k0);
}

template <typename T, typename S, int NC, int V, bool UPD, bool CC, bool NT, int UCAP,
          bool UA = false, bool NTA = NT>
HPTTX_HD void row_body(const RowDesc &d, const RowCoord &rc, uint32_t tid,
                       const HPTTX_GLOBAL T *HPTTX_RESTRICT A,
                       HPTTX_GLOBAL T *HPTTX_RESTRICT B,
                       const int64_t *tabA, const int64_t *tabB, const Coef<S> &k) {
  if (!UA || rc.slen >= (uint32_t)V)
    row_pass<T, S, NC, V, UPD, CC, NT, UCAP, UA, true, NTA>(d, rc, tid, A, B, tabA, tabB, k);
  else
    row_pass<T, S, NC, V, UPD, CC, NT, UCAP, UA, false, NTA>(d, rc, tid, A, B, tabA, tabB, k);
}

}  // namespace hpttx
