// planner.cpp -- see planner.h.
#include "planner.h"

#include <algorithm>
#include <cstdio>
#include <sstream>

namespace hpttx {

uint32_t dtype_esize(int dtype) {
  switch (dtype) {
    case kF32: return 4;
    case kF64: return 8;
    case kC64: return 8;
    case kC128: return 16;
    default: return 0;
  }
}

int build_problem(int dtype, uint32_t rank, const uint32_t *in_size,
                  const uint32_t *perm, const uint32_t *in_outer,
                  const uint32_t *out_outer, Problem &p, std::string &err) {
  p = Problem();
  p.dtype = dtype;
  p.esize = dtype_esize(dtype);
  if (p.esize == 0) { err = "unknown dtype"; return kEInval; }
  // src/hptt.cc:30-33: order <= 1 -> nullptr.  Ranks above 6 are rejected
  // rather than accepted and silently ignored (hptt_trans_gen.py:41-59).
  if (rank <= 1) { err = "rank must be at least 2"; return kEInval; }
  if (rank > (uint32_t)kMaxRank) { err = "rank above 6 is not supported"; return kEUnsupported; }
  if (!in_size || !perm) { err = "null size or perm array"; return kEInval; }
  // src/hptt.cc:49-56: permutation must be a bijection of [0, rank).
  bool seen[kMaxRank] = {false};
  for (uint32_t d = 0; d < rank; ++d) {
    if (perm[d] >= rank || seen[perm[d]]) { err = "perm is not a permutation"; return kEInval; }
    seen[perm[d]] = true;
  }
  // src/hptt.cc:41-47: zero sizes and outer < inner.  The output check uses
  // out_outer[d] >= in_size[perm[d]] (B dim d holds A dim perm[d]).
  for (uint32_t d = 0; d < rank; ++d) {
    if (in_size[d] == 0) { err = "zero-sized dim"; return kEInval; }
    if (in_outer && in_outer[d] < in_size[d]) { err = "in_outer_size smaller than in_size"; return kEInval; }
    if (out_outer && out_outer[d] < in_size[perm[d]]) { err = "out_outer_size smaller than the output size"; return kEInval; }
  }
  p.rank = rank;
  uint32_t out_size[kMaxRank];
  for (uint32_t d = 0; d < rank; ++d) {
    p.in_size[d] = in_size[d];
    p.perm[d] = perm[d];
    p.in_outer[d] = in_outer ? in_outer[d] : in_size[d];
    out_size[d] = in_size[perm[d]];
    p.out_outer[d] = out_outer ? out_outer[d] : out_size[d];
  }
  // Column-major strides over the outer sizes (inc/hptt/tensor.tcc:186-196).
  int64_t sa[kMaxRank], sbB[kMaxRank];
  sa[0] = 1;
  sbB[0] = 1;
  for (uint32_t d = 0; d + 1 < rank; ++d) {
    sa[d + 1] = sa[d] * (int64_t)p.in_outer[d];
    sbB[d + 1] = sbB[d] * (int64_t)p.out_outer[d];
  }
  uint64_t spanA = 1, spanB = 1, N = 1;
  for (uint32_t d = 0; d < rank; ++d) {
    spanA += (uint64_t)(in_size[d] - 1) * (uint64_t)sa[d];
    spanB += (uint64_t)(out_size[d] - 1) * (uint64_t)sbB[d];
    N *= in_size[d];
  }
  p.spanA = spanA;
  p.spanB = spanB;
  p.N = N;
  // A dim i sits at B dim q with perm[q] = i.
  int64_t sb_of_a[kMaxRank];
  for (uint32_t q = 0; q < rank; ++q) sb_of_a[perm[q]] = sbB[q];
  // Squeeze size-1 dims, then fuse neighbours that are contiguous in both A
  // and B.  This is the reference's merge rule (parameter_trans.tcc:235-250:
  // perm[k] == perm[k-1]+1 with unpadded inner dims) stated on strides.
  int n = 0;
  FusedDim tmp[kMaxRank];
  for (uint32_t i = 0; i < rank; ++i) {
    if (in_size[i] == 1) continue;
    tmp[n++] = FusedDim{in_size[i], sa[i], sb_of_a[i]};
  }
  int m = 0;
  for (int i = 0; i < n; ++i) {
    if (m > 0) {
      FusedDim &last = p.d[m - 1];
      if (tmp[i].sa == last.sa * (int64_t)last.n && tmp[i].sb == last.sb * (int64_t)last.n) {
        last.n *= tmp[i].n;
        continue;
      }
    }
    p.d[m++] = tmp[i];
  }
  // Kernels index a fused dim with 32 bits: split any dim of 2^31 elements
  // or more into (n / a, a), a its smallest divisor that leaves both parts
  // below 2^31, searched among the 2^20 candidates above n / 2^31 (an 8 GiB
  // fp32 tensor fuses to one such dim under an identity perm; 288 GB of HBM
  // holds tensors far larger).
  for (int i = 0; i < m; ++i) {
    const uint64_t n = p.d[i].n;
    if (n < (1ull << 31)) continue;
    if (m >= kMaxRank) {
      err = "a fused dim of 2^31 or more elements in a rank-6 problem";
      return kEUnsupported;
    }
    uint64_t a = (n >> 31) + 1;
    const uint64_t a_end = std::min<uint64_t>(1ull << 31, a + (1ull << 20));   // bounded search
    while (a < a_end && n % a) ++a;
    if (a >= a_end || n / a >= (1ull << 31)) {
      err = "a fused dim cannot be split below 2^31";
      return kEUnsupported;
    }
    for (int k = m; k > i + 1; --k) p.d[k] = p.d[k - 1];
    const FusedDim whole = p.d[i];
    p.d[i] = FusedDim{n / a, whole.sa, whole.sb};
    p.d[i + 1] = FusedDim{a, whole.sa * (int64_t)(n / a), whole.sb * (int64_t)(n / a)};
    ++m;
    --i;   // re-check the first part (already below 2^31)
  }
  p.fr = m;
  for (int i = 0; i < m; ++i) p.border[i] = i;
  std::sort(p.border, p.border + m, [&](int x, int y) { return p.d[x].sb < p.d[y].sb; });
  return kOk;
}

uint32_t align_vec(const void *ptr, uint32_t esize, uint32_t cap) {
  uintptr_t a = (uintptr_t)ptr;
  uint32_t v = cap;
  while (v > 1 && (a % (uintptr_t)(v * esize)) != 0) v >>= 1;
  return v;
}

static void set_dimlist(DimList &L) {
  // caller filled nd, div[i].d (as plain extents), sa, sb
  for (uint32_t i = 0; i < L.nd; ++i) L.div[i] = make_fastdiv(L.div[i].d);
  for (uint32_t i = L.nd; i < (uint32_t)kMaxRank; ++i) {
    L.div[i] = make_fastdiv(1);
    L.sa[i] = 0;
    L.sb[i] = 0;
  }
}

static bool divisible(int64_t x, uint32_t v) { return (x % (int64_t)v) == 0; }

TuneParams default_tune(const Problem &p) {
  TuneParams t;
  // Read runs of 256 bytes along A, write runs of 512 bytes along B (1024
  // bytes for 8-byte elements), and tiles of at least ~4 KiB: what exhaustive
  // tuning of the 6-D fp32 / fp64 and 4-D complex<double> workloads picked on
  // MI355X (64x128, 32x128 and 16x32 tiles).  The fp64 32x128 tile (34 KiB,
  // 512-thread blocks) measured 3.5% faster than 32x64 on the 6-D workload
  // and 1% faster on the batched one.
  t.run_r = 256 / p.esize;
  t.run_c = (p.esize == 8 ? 1024 : 512) / p.esize;
  t.block_elems = 8192 / p.esize;
  t.wide_min_bytes = 128;
  return t;
}

static void fill_tune_defaults(const Problem &p, TuneParams &t) {
  TuneParams d = default_tune(p);
  if (!t.run_r) t.run_r = d.run_r;
  if (!t.run_c) t.run_c = d.run_c;
  if (!t.block_elems) t.block_elems = d.block_elems;
  if (!t.wide_min_bytes) t.wide_min_bytes = d.wide_min_bytes;
}

// ROWS plan over a row dim `rowdim` (index into p.d, or -1 for a single
// element) and every other dim as the row index walked in B order.
static int plan_rows(const Problem &p, int rowdim, const TuneParams &tp,
                     LaunchPlan &lp, std::string &err) {
  RowDesc &r = lp.rows;
  r = RowDesc();
  const uint32_t vcap = 16 / p.esize;
  uint64_t L = rowdim >= 0 ? p.d[rowdim].n : 1;
  r.ra = rowdim >= 0 ? p.d[rowdim].sa : 1;
  r.rb = rowdim >= 0 ? p.d[rowdim].sb : 1;
  uint64_t nrows = 1;
  DimList &W = r.W;
  W.nd = 0;
  for (int k = 0; k < p.fr; ++k) {
    int i = p.border[k];
    if (i == rowdim) continue;
    W.div[W.nd].d = (uint32_t)p.d[i].n;
    W.sa[W.nd] = p.d[i].sa;
    W.sb[W.nd] = p.d[i].sb;
    W.nd++;
    nrows *= p.d[i].n;
  }
  if (L > 0xFFFFFFFFull || nrows > 0xFFFFFFFFull) {
    err = "row extent or row count exceeds 2^32";
    return kEUnsupported;
  }
  set_dimlist(W);
  r.L = (uint32_t)L;
  r.nrows = (uint32_t)nrows;
  // Vector width the shape allows: contiguous rows whose length and every
  // row offset are multiples of v.
  uint32_t v = 1;
  if (r.ra == 1 && r.rb == 1) {
    v = vcap;
    while (v > 1) {
      bool okv = (L % v) == 0;
      for (uint32_t i = 0; i < W.nd && okv; ++i)
        okv = divisible(W.sa[i], v) && divisible(W.sb[i], v);
      if (okv) break;
      v >>= 1;
    }
  }
  lp.vr_shape = lp.vw_shape = v;
  // A block tile covers ~block_elems elements: one segment of RB rows.
  // Rows longer than that are cut into equal segments (a multiple of a
  // 128-byte line where that keeps the segment count; else of the vector):
  // 2144-element fp32 rows went from segments of 2048 + 96 to 1088 + 1056.
  const uint32_t be = std::max<uint32_t>(tp.block_elems, kBlock);
  uint32_t seg = (uint32_t)std::min<uint64_t>(L, be);
  if (seg < L) {
    const uint64_t nseg = (L + be - 1) / be;
    const uint64_t even = (L + nseg - 1) / nseg;
    const uint64_t line = std::max<uint64_t>(vcap, 128 / p.esize);
    uint64_t s2 = (even + line - 1) / line * line;
    if ((L + s2 - 1) / s2 != nseg) s2 = (even + vcap - 1) / vcap * vcap;
    seg = (uint32_t)std::max<uint64_t>(v, std::min<uint64_t>(s2, L));
  }
  uint32_t rb = std::max<uint32_t>(1, be / std::max<uint32_t>(seg, 1));
  rb = std::min<uint32_t>(rb, 1024);
  rb = (uint32_t)std::min<uint64_t>(rb, nrows);
  r.SEG = seg;
  r.nseg = (uint32_t)((L + seg - 1) / seg);
  r.RB = rb;
  r.nrb = (uint32_t)((nrows + rb - 1) / rb);
  uint64_t nt = (uint64_t)r.nseg * r.nrb;
  if (nt > 0xFFFFFFFFull) { err = "too many tiles"; return kEUnsupported; }
  r.ntiles = (uint32_t)nt;
  r.div_nseg = make_fastdiv(r.nseg);
  lp.kind = kRows;
  lp.tune = tp;
  if (rowdim < 0 || p.fr == 1)
    lp.path = "axpby";
  else if (r.ra == 1 && r.rb == 1)
    lp.path = "wide-copy";
  else
    lp.path = "gather";
  return kOk;
}

static int plan_tile(const Problem &p, bool shared, const TuneParams &tp, uint32_t min_tiles,
                     LaunchPlan &lp, std::string &err) {
  TileDesc &t = lp.tile;
  t = TileDesc();
  const uint32_t vcap = 16 / p.esize;
  const int a0 = 0;              // A's stride-1 dim
  const int b0 = p.border[0];    // B's stride-1 dim
  const uint32_t VS = shared ? (uint32_t)p.d[0].n : 1;
  bool inR[kMaxRank] = {false}, inC[kMaxRank] = {false};
  // R grows along A order from the first dim after S, C along B order.
  int ra = shared ? 1 : a0;      // index into A order
  int cb = shared ? 1 : 0;       // index into B order
  if (ra >= p.fr || cb >= p.fr) { err = "internal: tile groups empty"; return kEInval; }
  if (inC[ra]) { err = "internal"; return kEInval; }
  std::vector<int> Rd{ra}, Cd{p.border[cb]};
  inR[ra] = true;
  inC[p.border[cb]] = true;
  if (Rd[0] == Cd[0]) { err = "internal: R and C share a dim"; return kEInval; }
  // contiguity of the first R dim after S in A, first C dim after S in B
  if (shared) {
    if (p.d[ra].sa != (int64_t)VS || p.d[p.border[cb]].sb != (int64_t)VS) {
      err = "internal: shared group not contiguous";
      return kEInval;
    }
  } else {
    if (p.d[a0].sa != 1 || p.d[b0].sb != 1) { err = "internal: stride-1 dims"; return kEInval; }
  }
  uint64_t VR = p.d[ra].n, VC = p.d[p.border[cb]].n;
  for (bool grew = true; grew;) {
    grew = false;
    if (VR * VS < tp.run_r && ra + 1 < p.fr) {
      int nx = ra + 1;
      const FusedDim &cur = p.d[ra];
      if (!inC[nx] && p.d[nx].sa == cur.sa * (int64_t)cur.n) {
        ra = nx;
        inR[nx] = true;
        Rd.push_back(nx);
        VR *= p.d[nx].n;
        grew = true;
      }
    }
    if (VC * VS < tp.run_c && cb + 1 < p.fr) {
      int nx = p.border[cb + 1];
      const FusedDim &cur = p.d[p.border[cb]];
      if (!inR[nx] && p.d[nx].sb == cur.sb * (int64_t)cur.n) {
        cb = cb + 1;
        inC[nx] = true;
        Cd.push_back(nx);
        VC *= p.d[nx].n;
        grew = true;
      }
    }
  }
  std::vector<int> Od;
  uint64_t VO = 1;
  for (int i = 0; i < p.fr; ++i) {
    if (inR[i] || inC[i] || (shared && i == 0)) continue;
    Od.push_back(i);
    VO *= p.d[i].n;
  }
  if (VR > 0xFFFFFFFFull || VC > 0xFFFFFFFFull || VO > 0xFFFFFFFFull) {
    err = "a tile super-dim exceeds 2^32 indices";
    return kEUnsupported;
  }
  t.VS = VS;
  t.VR = (uint32_t)VR;
  t.VC = (uint32_t)VC;
  t.VO = (uint32_t)VO;
  auto fill = [&](DimList &L, const std::vector<int> &dims) {
    L.nd = (uint32_t)dims.size();
    for (size_t k = 0; k < dims.size(); ++k) {
      L.div[k].d = (uint32_t)p.d[dims[k]].n;
      L.sa[k] = p.d[dims[k]].sa;
      L.sb[k] = p.d[dims[k]].sb;
    }
    set_dimlist(L);
  };
  fill(t.R, Rd);
  fill(t.C, Cd);
  fill(t.O, Od);
  // Vector widths the shape allows.  Reads: runs r0*VS + [0, rlen*VS) at A
  // offsets from C and O; writes: runs at B offsets from R and O.
  uint32_t vr = vcap, vw = vcap;
  auto ok_r = [&](uint32_t v) {
    if (((uint64_t)VS * VR) % v) return false;
    for (int i : Cd) if (!divisible(p.d[i].sa, v)) return false;
    for (int i : Od) if (!divisible(p.d[i].sa, v)) return false;
    return true;
  };
  auto ok_w = [&](uint32_t v) {
    if (((uint64_t)VS * VC) % v) return false;
    for (int i : Rd) if (!divisible(p.d[i].sb, v)) return false;
    for (int i : Od) if (!divisible(p.d[i].sb, v)) return false;
    return true;
  };
  while (vr > 1 && !ok_r(vr)) vr >>= 1;
  while (vw > 1 && !ok_w(vw)) vw >>= 1;
  // Tile extents.  Start from runs of ~run_r / run_c elements along A / B,
  // then, while the tile holds less than half of block_elems, lengthen the
  // shorter run (doubling) as far as its super-dim allows; the O count fills
  // the rest.  Each extent is then balanced over its tile count (no sliver
  // edge tiles) and rounded to whole 16-byte vectors (interior tiles then
  // need no partial vector even where the runs are unaligned).
  auto cdiv = [](uint64_t a, uint64_t b) { return (a + b - 1) / b; };
  const bool beyond_llc = p.N * p.esize * 2 > kLastLevelCacheBytes;
  auto balance = [&](uint64_t V, uint64_t T, uint32_t vec, bool is_r) {
    const uint64_t n = cdiv(V, T);
    uint64_t b = cdiv(V, n);
    // vector granularity in index units: runs are index*VS elements
    uint64_t g = 1;
    while (g < vec && ((g * VS) % vec) != 0) ++g;
    if (g > 1) b = std::min<uint64_t>(V, cdiv(b, g) * g);
    // Problems larger than the last-level cache: runs of whole 128-byte
    // lines where that keeps the tile within the requested run and the tile
    // count unchanged (the nearest multiple of the line, else the next one
    // up).  A run that ends inside a line shares it with the neighbouring
    // tile, which another XCD (another L2) fetches from HBM again once the
    // line has left the last-level cache.  Measured on the reference's
    // 102-shape list: median +5%, up to +47% (tiles of 60 x 124 fp32 become
    // 64 x 128).  Smaller problems, whose shared lines the last-level cache
    // serves, and shared-group tiles measured slower with it.
    if (VS == 1 && V > b && beyond_llc) {
      const uint64_t line = 128 / p.esize;
      const uint64_t cap_r = std::min<uint64_t>(V, std::max<uint64_t>(T, b));
      const uint64_t lo = std::max<uint64_t>(line, (b + line / 2) / line * line);
      const uint64_t hi = cdiv(b, line) * line;
      if (lo <= cap_r && cdiv(V, lo) == n) b = lo;
      else if (hi <= cap_r && cdiv(V, hi) == n) b = hi;
      // An R extent of at most 4 tiles that the tile does not divide (96
      // fp32 elements as 64 + 32): equal runs of whole lines that divide it
      // (3 x 32) where there are such, so that no block gets a short last
      // tile.  Measured on MI355X: the reference's 27 shapes with R = 96 and
      // 4 with R = 160 (fp32) +0.1% to +8.3%, 5% on average; 8 fp64, c64 and
      // c128 reversals and perms with R of 1.5 to 2.5 tiles (48 and 80 fp64
      // as 16s, 24 and 40 c128 as 8s) +4% to +12%; longer extents
      // (352 to 2400, 5.5 to 37.5 tiles) -11% to +7%, and the same rule on C
      // +1% at most, so neither is split.
      if (is_r && V % b != 0 && cdiv(V, b) <= 4) {
        for (uint64_t d = b / line * line; d >= line; d -= line)
          if (V % d == 0) {
            b = d;
            break;
          }
      }
    }
    return std::max<uint64_t>(1, b);
  };
  uint64_t TR = 1, TC = 1, TO = 1;
  auto size_tile = [&](uint32_t run_r, uint32_t run_c, uint32_t block) {
    TR = std::min<uint64_t>(VR, std::max<uint64_t>(1, cdiv(run_r, VS)));
    TC = std::min<uint64_t>(VC, std::max<uint64_t>(1, cdiv(run_c, VS)));
    const uint64_t half_block = std::max<uint64_t>(1, block / 2);
    while (TR * TC * VS < half_block) {
      const bool can_r = TR < VR, can_c = TC < VC;
      if (!can_r && !can_c) break;
      const bool grow_r = can_r && (!can_c || TR <= TC);
      if (grow_r) TR = std::min<uint64_t>(VR, TR * 2);
      else TC = std::min<uint64_t>(VC, TC * 2);
    }
    TR = balance(VR, TR, vcap, true);
    TC = balance(VC, TC, vcap, false);
    const uint64_t per_o = TR * VS * TC;
    TO = std::max<uint64_t>(1, block / std::max<uint64_t>(per_o, 1));
    TO = std::min<uint64_t>(std::min<uint64_t>(TO, VO), 512);
    TO = cdiv(VO, cdiv(VO, TO));
    return cdiv(VR, TR) * cdiv(VC, TC) * cdiv(VO, TO);
  };
  uint32_t run_r = tp.run_r, run_c = tp.run_c, block = tp.block_elems;
  uint64_t count = size_tile(run_r, run_c, block);
  // Small problems: halve the runs (read runs down to 128 bytes, write runs
  // down to 256) until the launch has min_tiles blocks to spread over the
  // CUs.
  while (count < min_tiles) {
    const uint64_t rb = (uint64_t)run_r * p.esize, cb = (uint64_t)run_c * p.esize;
    const bool can_r = rb > 128 && TR * VS > 1, can_c = cb > 256 && TC * VS > 1;
    if (!can_r && !can_c) break;
    if (can_r && (!can_c || rb >= cb / 2)) run_r /= 2;
    else run_c /= 2;
    block = std::max<uint32_t>(kBlock, block / 2);
    count = size_tile(run_r, run_c, block);
  }
  // Whole-tile runs must stay multiples of the exact vector widths.
  while (vr > 1 && (TR * VS) % vr) vr >>= 1;
  while (vw > 1 && (TC * VS) % vw) vw >>= 1;
  t.TR = (uint32_t)TR;
  t.TC = (uint32_t)TC;
  t.TO = (uint32_t)TO;
  t.nR = (uint32_t)((VR + TR - 1) / TR);
  t.nC = (uint32_t)((VC + TC - 1) / TC);
  t.nO = (uint32_t)((VO + TO - 1) / TO);
  uint64_t nt = (uint64_t)t.nR * t.nC * t.nO;
  if (nt > 0xFFFFFFFFull) { err = "too many tiles"; return kEUnsupported; }
  t.ntiles = (uint32_t)nt;
  t.div_nR = make_fastdiv(t.nR);
  t.div_nC = make_fastdiv(t.nC);
  t.div_TC = make_fastdiv(TC);
  t.div_TR = make_fastdiv(TR);
  t.div_VS = make_fastdiv(VS);
  lp.vr_shape = vr;
  lp.vw_shape = vw;
  lp.kind = kTile;
  lp.tune = tp;
  lp.path = shared ? "tile-shared" : "tile";
  return kOk;
}

void finalize_vectors(const Problem &p, LaunchPlan &lp, uint32_t vr, uint32_t vw) {
  // vr / vw: the full width (16 / esize) where runs are exact (aligned, whole
  // vectors); anything less selects unaligned full-width vectors with a
  // partial last vector, or scalars for 16-byte elements.
  const uint32_t cap = 16 / p.esize;
  if (lp.kind == kTile) {
    TileDesc &t = lp.tile;
    t.ua_r = (cap > 1 && vr != cap) ? 1u : 0u;
    t.ua_w = (cap > 1 && vw != cap) ? 1u : 0u;
    t.vr = cap;
    t.vw = cap;
    t.nvr = (t.TR * t.VS + cap - 1) / cap;
    t.nvw = (t.TC * t.VS + cap - 1) / cap;
    t.div_nvr = make_fastdiv(t.nvr);
    t.div_nvw = make_fastdiv(t.nvw);
    t.div_ngw = make_fastdiv((t.TR + cap - 1) / cap);
    // LDS rows: a multiple of the read vector.  Rows of a power-of-two
    // number (>= 4) of vectors use the chunk swizzle (measured 1.4% faster
    // than padding on the 6-D fp32 workload); others are padded by one
    // vector (one element for scalars, to an odd stride) so that the write
    // pass's column gathers spread over banks.
    uint32_t row = t.TR * t.VS;
    row = (row + cap - 1) / cap * cap;
    t.lds_row = row + cap;
    if (cap == 1 && (row % 2) == 0) t.lds_row = row + 1;
    t.swz_mask = 0;
    const uint32_t chunks = row / cap;
    if (cap > 1 && t.VS == 1 && chunks >= 4 && (chunks & (chunks - 1)) == 0 && t.TC % cap == 0) {
      t.lds_row = row;
      t.swz_mask = chunks - 1;
    }
    // Shared groups of whole vectors: the write pass reads whole vectors.
    t.sh_vec = (cap > 1 && t.VS > 1 && t.VS % cap == 0) ? 1u : 0u;
    // XCD-aware tile order (each XCD runs 8 consecutive tiles) in problems
    // beyond the last-level cache: neighbouring tiles read and write
    // neighbouring pieces of the same rows, and one XCD then fetches a row's
    // lines and DRAM pages together instead of eight XCDs fetching them at
    // different times.  Measured on MI355X against the plain order: the
    // reference's 102 workloads +1.7% (geometric mean; -1.9% to +10.9%),
    // c128 {3,2,1,0} reversals of {64,128,128,64} +46%, {97,97,97,97} +11%,
    // {100,100,100,100} +6%, {96,96,96,96} unchanged, 6-D fp32 +0.9%, 6-D
    // fp64 -1.4%.  Runs of 16 or 32 consecutive tiles per XCD were no better.
    t.xcd = p.N * p.esize * 2 > kLastLevelCacheBytes ? 1u : 0u;
    lp.lds_bytes = tile_lds_bytes(t, p.esize);
    lp.threads = lp.lds_bytes > kWideBlockMinLds ? kWideBlock : kBlock;
    lp.grid = t.ntiles;
  } else {
    RowDesc &r = lp.rows;
    const bool contiguous = r.ra == 1 && r.rb == 1;
    r.v = contiguous ? cap : 1;
    r.ua = (contiguous && cap > 1 && vr != cap) ? 1u : 0u;
    r.nvs = (r.SEG + r.v - 1) / r.v;
    r.div_nvs = make_fastdiv(r.nvs);
    lp.lds_bytes = row_lds_bytes(r);
    lp.threads = kBlock;
    lp.grid = r.ntiles;
  }
  (void)vw;
}

// Every run length and extent, and every stride that places a run start,
// a multiple of `unit` bytes.
static bool tile_runs_on(const Problem &p, const TileDesc &t, uint64_t unit) {
  auto lines = [&](uint64_t elems) { return (elems * p.esize) % unit == 0; };
  if (!lines((uint64_t)t.TR * t.VS) || !lines((uint64_t)t.VR * t.VS) ||
      !lines((uint64_t)t.TC * t.VS) || !lines((uint64_t)t.VC * t.VS))
    return false;
  for (uint32_t k = 0; k < t.C.nd; ++k)
    if (!lines((uint64_t)t.C.sa[k])) return false;
  for (uint32_t k = 0; k < t.R.nd; ++k)
    if (!lines((uint64_t)t.R.sb[k])) return false;
  for (uint32_t k = 0; k < t.O.nd; ++k)
    if (!lines((uint64_t)t.O.sa[k]) || !lines((uint64_t)t.O.sb[k])) return false;
  return true;
}

bool tile_whole_lines(const Problem &p, const TileDesc &t) { return tile_runs_on(p, t, 128); }

// Each buffer at most the last-level cache's size (A and B together are
// larger, or nothing streams).
static bool buffers_fit_cache(const Problem &p) {
  return std::max(p.spanA, p.spanB) * p.esize <= kLastLevelCacheBytes;
}

// Non-temporal choice of a tile plan whose runs are not all whole lines:
// buffers that fit the cache stream loads only (4-byte elements) or nothing;
// larger ones stream everything for 4-byte elements whose runs all start and
// end on 64-byte half lines, and nothing otherwise.
static void stream_partial_lines(const Problem &p, LaunchPlan &lp, bool half_lines) {
  if (buffers_fit_cache(p)) {
    lp.nt_load = lp.nt_load_b = p.esize == 4;
    lp.nt_store = false;
  } else {
    lp.nt_load = lp.nt_load_b = lp.nt_store = p.esize == 4 && half_lines;
  }
}

void stream_for_pointers(const Problem &p, LaunchPlan &lp, const void *A, const void *B) {
  const uintptr_t a = (uintptr_t)A | (uintptr_t)B;
  if (lp.kind == kTile && lp.nt_store && lp.tune.stream == 0 && a % 128 != 0)
    stream_partial_lines(p, lp, a % 64 == 0 && tile_runs_on(p, lp.tile, 64));
}

int plan_launch(const Problem &p, const TuneParams &tp_in, LaunchPlan &lp,
                std::string &err, uint32_t min_tiles) {
  TuneParams tp = tp_in;
  fill_tune_defaults(p, tp);
  lp = LaunchPlan();
  int rc;
  if (p.fr == 0) {
    rc = plan_rows(p, -1, tp, lp, err);            // a single element
  } else if (p.fr == 1) {
    rc = plan_rows(p, 0, tp, lp, err);             // identity: axpby
  } else if (p.d[0].sa != 1 || p.d[p.border[0]].sb != 1) {
    rc = plan_rows(p, p.border[0], tp, lp, err);   // no unit-stride dim (padded size-1 dims)
  } else if (p.border[0] == 0) {
    // Stride-1 dim kept (reference "common leading",
    // parameter_trans.tcc:154-158).  Rows of at least wide_min_bytes (128 by
    // default): wide copy; shorter rows: TILE with the rows as a shared
    // innermost group.
    const uint64_t L = p.d[0].n;
    const bool contiguous_after = p.d[1].sa == (int64_t)L && p.d[p.border[1]].sb == (int64_t)L;
    if (L * p.esize >= tp.wide_min_bytes || !contiguous_after || p.border[1] == 1)
      rc = plan_rows(p, 0, tp, lp, err);
    else
      rc = plan_tile(p, true, tp, min_tiles, lp, err);
  } else {
    rc = plan_tile(p, false, tp, min_tiles, lp, err);
  }
  if (rc != kOk) return rc;
  // Streaming mode.  Model: on when A and B together exceed the 256 MiB
  // last-level cache and every access belongs to a run of at least 128
  // bytes.  Rows (the wide copy) stream every access (of the reference's
  // shapes with kept rows of 128-255 bytes: 192-byte rows 2.9 -> 5.3 TB/s
  // against a shared-group tile), except A's loads where each buffer fits
  // in the cache (the reference's 12 such row shapes +11% to +20% back to
  // back, -1.6% to +5% with a cold cache).  Tiles:
  //  * runs of whole 128-byte lines: every access streamed; where each
  //    buffer fits in the cache, A's loads are not, so that A stays there
  //    for the next call: measured on MI355X, calls repeated back to back,
  //    the reference's 57 such fp32 shapes (about 200 MB per buffer) +7.5%
  //    to +30% (23% on average), fp64, c64 and c128 reversals of 198 MB
  //    +17% to +22%, and with a cold cache (a 512 MB read before each call)
  //    -0.9% to +0.5%; at 2 to 4 times those sizes the same
  //    choice lost 1%, and the 6-D {16,24,32,16,24,32} perm {5,2,0,4,1,3}
  //    (604 MB) 1.2% to 1.7%, so larger buffers stream everything;
  //  * runs that are not all whole lines (stream_partial_lines; the
  //    pointers' part is checked when the plan gets its data,
  //    stream_for_pointers): buffers that fit, loads only (4-byte elements;
  //    back to back their stores keep B in the cache) or nothing; larger
  //    ones, everything for 4-byte elements on 64-byte half lines and
  //    nothing otherwise: of 10 fp32 shapes of 660 MB to 2 GB per buffer,
  //    +4% to +27% over loads only on every one (everything streamed best on
  //    3 of 4 on half lines, the fourth 0.8% behind nothing; nothing best on
  //    5 of 6 off them, the sixth 2.5% behind everything).
  // The tuner may override.
  const bool rows = lp.kind == kRows;
  if (tp.stream == 1 || (rows && (tp.stream == 3 || tp.stream == 4))) {
    lp.nt_load = lp.nt_load_b = lp.nt_store = true;
  } else if (tp.stream == 2) {
    lp.nt_load = lp.nt_load_b = lp.nt_store = false;
  } else if (tp.stream == 3) {
    lp.nt_load = lp.nt_load_b = true;
    lp.nt_store = false;
  } else if (tp.stream == 4) {
    lp.nt_load = lp.nt_load_b = false;
    lp.nt_store = true;
  } else if (tp.stream == 5) {
    lp.nt_load = false;
    lp.nt_load_b = lp.nt_store = true;
  } else {
    const uint64_t bytes = p.N * p.esize * 2;
    bool long_runs;
    if (!rows)
      long_runs = (uint64_t)lp.tile.TR * lp.tile.VS * p.esize >= 128 &&
                  (uint64_t)lp.tile.TC * lp.tile.VS * p.esize >= 128;
    else
      long_runs = lp.rows.ra == 1 && lp.rows.rb == 1 && (uint64_t)lp.rows.L * p.esize >= 128;
    lp.nt_load = lp.nt_load_b = lp.nt_store = long_runs && bytes > kLastLevelCacheBytes;
    if (lp.nt_load) {
      if (!rows && !tile_whole_lines(p, lp.tile))
        stream_partial_lines(p, lp, tile_runs_on(p, lp.tile, 64));
      else if (buffers_fit_cache(p))
        lp.nt_load = false;   // A stays in the cache for the next call (tune stream 5)
    }
  }
  finalize_vectors(p, lp, lp.vr_shape, lp.vw_shape);
  if (lp.lds_bytes > 65536) { err = "tile needs more than 64 KiB of LDS"; return kEUnsupported; }
  return kOk;
}

std::string describe(const Problem &p, const LaunchPlan &lp) {
  std::ostringstream os;
  static const char *names[] = {"float", "double", "complex<float>", "complex<double>"};
  os << "dtype: " << names[p.dtype] << "  rank: " << p.rank << "  elements: " << p.N << "\n";
  os << "fused dims (A order, size:strideA:strideB):";
  for (int i = 0; i < p.fr; ++i) os << " " << p.d[i].n << ":" << p.d[i].sa << ":" << p.d[i].sb;
  os << "\nB order:";
  for (int i = 0; i < p.fr; ++i) os << " " << p.border[i];
  os << "\npath: " << lp.path;
  if (lp.kind == kTile) {
    const TileDesc &t = lp.tile;
    os << "\ntile: VS=" << t.VS << " R=" << t.VR << "/" << t.TR << " C=" << t.VC << "/" << t.TC
       << " O=" << t.VO << "/" << t.TO << " vec(r,w)=" << t.vr << (t.ua_r ? "u" : "") << ","
       << t.vw << (t.ua_w ? "u" : "")
       << " tiles=" << t.ntiles << " lds=" << lp.lds_bytes << "B"
       << (t.swz_mask ? " swizzled" : " padded")
       << (t.xcd ? " xcd-order" : "") << (t.sh_vec ? " shared-vectors" : "");
  } else {
    const RowDesc &r = lp.rows;
    os << "\nrows: L=" << r.L << " stride(a,b)=" << r.ra << "," << r.rb << " rows=" << r.nrows
       << " seg=" << r.SEG << " rows/block=" << r.RB << " vec=" << r.v << (r.ua ? "u" : "")
       << " tiles=" << r.ntiles;
  }
  os << "\ngrid: " << lp.grid << " x " << lp.threads << "  streaming: "
     << (lp.nt_load ? (lp.nt_store ? "on" : "loads")
                    : (lp.nt_store ? (lp.nt_load_b ? "stores+b-loads" : "stores") : "off"))
     << "  tune(run_r,run_c,block,wide,stream)=" << lp.tune.run_r << "," << lp.tune.run_c << ","
     << lp.tune.block_elems << "," << lp.tune.wide_min_bytes << "," << lp.tune.stream << "\n";
  return os.str();
}

}  // namespace hpttx
