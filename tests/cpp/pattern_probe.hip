// pattern_probe.hip -- TEST INFRASTRUCTURE: the rate of a transposition's
// memory access pattern on the GPU it runs on, without the transposition
// (never part of the product).
//
// For a workload's tile decomposition (R: dims contiguous in A, C: dims
// contiguous in B, O: the rest; TR x TC tiles, one tile per block, tiles
// numbered R fastest, then C, then O, as the library's tile kernel numbers
// them) each block
//   1. loads its tile's A runs (TC runs of TR elements, 16-byte vectors),
//   2. waits at a barrier (mode "pattern"; mode "no-barrier" skips it),
//   3. loads and stores its tile's B runs (TR runs of TC elements),
//      storing alpha * (the vector it loaded from A) + beta * B.
// Values are not transposed (a thread stores what it loaded itself), so no
// LDS is used; addresses and bytes are exactly the tile kernel's.  Mode
// "sequential" runs the same skeleton (same threads, vectors per thread and
// tile count) over A and B in address order: the skeleton's streaming rate.
// Loads and stores are non-temporal, as the library's streaming mode issues
// them.  Every address is checked on the host before the first launch.
//
// GB/s = (read A + read B + write B) bytes / mean duration of 20
// back-to-back launches after 3 warm-up launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

typedef float f4 __attribute__((ext_vector_type(4)));

#define CK(x)                                                                        \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      std::fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
      std::exit(1);                                                                  \
    }                                                                                \
  } while (0)

struct Desc {
  uint32_t TR, TC, nR, nC, nO, ntiles;
  uint32_t es;      // element bytes
  uint32_t nv;      // 16-byte vectors per tile (per side)
  uint32_t vr, vw;  // vectors per read run / per write run
};

constexpr int kMaxU = 4;  // vectors per thread per side

// Tile id of launch slot i.  order 0: R fastest, then C, then O (the tile
// kernel's order); 1: C fastest, then R, then O; 2: O fastest.
__host__ __device__ inline uint32_t slot_tile(uint32_t i, const Desc &d, int order) {
  if (order == 1) {
    const uint32_t ct = i % d.nC, q = i / d.nC, rt = q % d.nR, o = q / d.nR;
    return (o * d.nC + ct) * d.nR + rt;
  }
  if (order == 2) {
    const uint32_t o = i % d.nO, q = i / d.nO;
    return o * d.nR * d.nC + q;
  }
  return i;
}

// Byte offsets of vector v of tile t: side 0 reads A, side 1 reads and
// writes B.
__host__ __device__ inline int64_t vec_offset(const Desc &d, uint32_t t, uint32_t v, int side,
                                              bool sequential, const int64_t *AC,
                                              const int64_t *BR, const int64_t *AO,
                                              const int64_t *BO) {
  if (sequential) return ((int64_t)t * d.nv + v) * 16;
  const uint32_t rt = t % d.nR, q = t / d.nR, ct = q % d.nC, o = q / d.nC;
  if (side == 0) {
    const uint32_t row = v / d.vr, k = v - row * d.vr;
    return (AC[ct * d.TC + row] + AO[o] + (int64_t)rt * d.TR) * d.es + 16 * (int64_t)k;
  }
  const uint32_t row = v / d.vw, k = v - row * d.vw;
  return (BR[rt * d.TR + row] + BO[o] + (int64_t)ct * d.TC) * d.es + 16 * (int64_t)k;
}

// MODE 0: barrier between the phases; 1: no barrier; 2: sequential.
template <int MODE, int NTH>
__global__ __launch_bounds__(NTH) void k_pattern(const char *__restrict__ A,
                                                 char *__restrict__ B,
                                                 const int64_t *__restrict__ AC,
                                                 const int64_t *__restrict__ BR,
                                                 const int64_t *__restrict__ AO,
                                                 const int64_t *__restrict__ BO, Desc d,
                                                 int order, float al, float be) {
  const uint32_t t = slot_tile(blockIdx.x, d, order);
  f4 x[kMaxU];
#pragma unroll
  for (int u = 0; u < kMaxU; ++u) {
    const uint32_t v = threadIdx.x + u * NTH;
    if (v < d.nv) {
      const int64_t off = vec_offset(d, t, v, 0, MODE == 2, AC, BR, AO, BO);
      x[u] = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(A + off));
    }
  }
  if (MODE == 0) __syncthreads();
#pragma unroll
  for (int u = 0; u < kMaxU; ++u) {
    const uint32_t v = threadIdx.x + u * NTH;
    if (v < d.nv) {
      f4 *p = reinterpret_cast<f4 *>(B + vec_offset(d, t, v, 1, MODE == 2, AC, BR, AO, BO));
      const f4 y = __builtin_nontemporal_load(p);
      __builtin_nontemporal_store(al * x[u] + be * y, p);
    }
  }
}

struct Workload {
  const char *name;
  uint32_t es;
  int rank;
  int64_t n[6];
  int perm[6];  // B dim q is A dim perm[q]
  int nr, rd[4], nc, cd[4], no, od[4];  // R dims (A order), C dims (B order), O dims
  uint32_t TR, TC, nth;
  uint32_t lds;  // the tile kernel's LDS reservation for this tile (bytes)
};

struct Tables {
  std::vector<int64_t> AC, BR, AO, BO;
  int64_t N = 0;
};

// Offsets (in elements) of every index of a group of dims, first dim fastest.
static std::vector<int64_t> group_offsets(const int64_t *n, const int64_t *stride, int k,
                                          const int *dims) {
  std::vector<int64_t> out(1, 0);
  for (int j = 0; j < k; ++j) {
    const size_t m = out.size();
    std::vector<int64_t> next(m * n[dims[j]]);
    for (int64_t i = 0; i < n[dims[j]]; ++i)
      for (size_t b = 0; b < m; ++b) next[i * m + b] = out[b] + i * stride[dims[j]];
    out.swap(next);
  }
  return out;
}

static Tables make_tables(const Workload &w) {
  int64_t sa[6], sb[6];
  sa[0] = 1;
  for (int i = 1; i < w.rank; ++i) sa[i] = sa[i - 1] * w.n[i - 1];
  int64_t s = 1;
  for (int q = 0; q < w.rank; ++q) {
    sb[w.perm[q]] = s;
    s *= w.n[w.perm[q]];
  }
  Tables t;
  t.N = s;
  t.AC = group_offsets(w.n, sa, w.nc, w.cd);
  t.BR = group_offsets(w.n, sb, w.nr, w.rd);
  t.AO = group_offsets(w.n, sa, w.no, w.od);
  t.BO = group_offsets(w.n, sb, w.no, w.od);
  return t;
}

int main(int argc, char **argv) {
  const bool quick = argc > 1;  // any argument: the bench workloads' patterns only
  const Workload wls[] = {
      // 6-D {16,24,32,16,24,32} perm {5,2,0,4,1,3}: R = A dims 0,1 (384),
      // C = A dims 5,2 (B's first two, 1024), O = A dims 3,4 (384)
      {"6-D fp32 64x128", 4, 6, {16, 24, 32, 16, 24, 32}, {5, 2, 0, 4, 1, 3},
       2, {0, 1}, 2, {5, 2}, 2, {3, 4}, 64, 128, 512, 34320},
      {"6-D fp64 32x128", 8, 6, {16, 24, 32, 16, 24, 32}, {5, 2, 0, 4, 1, 3},
       2, {0, 1}, 2, {5, 2}, 2, {3, 4}, 32, 128, 512, 34064},
      // 4-D complex<double> {96,96,96,96} perm {3,2,1,0}: R = A dim 0,
      // C = A dim 3, O = A dims 1,2
      {"4-D c128 16x32", 16, 4, {96, 96, 96, 96}, {3, 2, 1, 0},
       1, {0}, 1, {3}, 2, {1, 2}, 16, 32, 256, 9104},
      // other tile shapes of the 6-D fp32 workload (same bytes per tile)
      {"6-D fp32 32x256", 4, 6, {16, 24, 32, 16, 24, 32}, {5, 2, 0, 4, 1, 3},
       2, {0, 1}, 2, {5, 2}, 2, {3, 4}, 32, 256, 512, 35088},
      {"6-D fp32 128x64", 4, 6, {16, 24, 32, 16, 24, 32}, {5, 2, 0, 4, 1, 3},
       2, {0, 1}, 2, {5, 2}, 2, {3, 4}, 128, 64, 512, 34320},
      {"6-D fp32 16x512", 4, 6, {16, 24, 32, 16, 24, 32}, {5, 2, 0, 4, 1, 3},
       2, {0, 1}, 2, {5, 2}, 2, {3, 4}, 16, 512, 512, 37008},
  };
  const int nwl = quick ? 3 : (int)(sizeof(wls) / sizeof(wls[0]));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  for (int wi = 0; wi < nwl; ++wi) {
    const Workload &w = wls[wi];
    const Tables tb = make_tables(w);
    Desc d;
    d.TR = w.TR;
    d.TC = w.TC;
    d.es = w.es;
    const uint32_t VR = (uint32_t)tb.BR.size(), VC = (uint32_t)tb.AC.size(),
                   VO = (uint32_t)tb.AO.size();
    if (VR % w.TR || VC % w.TC || (w.TR * w.es) % 16 || (w.TC * w.es) % 16) {
      std::fprintf(stderr, "%s: tile does not divide the extents\n", w.name);
      return 2;
    }
    d.nR = VR / w.TR;
    d.nC = VC / w.TC;
    d.nO = VO;
    d.ntiles = d.nR * d.nC * d.nO;
    d.vr = w.TR * w.es / 16;
    d.vw = w.TC * w.es / 16;
    d.nv = w.TR * w.TC * w.es / 16;
    if (d.nv > (uint32_t)kMaxU * w.nth || (uint64_t)d.ntiles * d.nv * 16 != (uint64_t)tb.N * w.es) {
      std::fprintf(stderr, "%s: tile does not fit the probe\n", w.name);
      return 2;
    }
    // every address of every mode and order, on the host
    const int64_t bytes = tb.N * w.es;
    for (uint32_t t = 0; t < d.ntiles; ++t)
      for (uint32_t v = 0; v < d.nv; ++v)
        for (int side = 0; side < 2; ++side) {
          const int64_t off = vec_offset(d, t, v, side, false, tb.AC.data(), tb.BR.data(),
                                         tb.AO.data(), tb.BO.data());
          if (off < 0 || off + 16 > bytes || off % 16) {
            std::fprintf(stderr, "%s: address out of range\n", w.name);
            return 3;
          }
        }
    char *A, *B;
    int64_t *dAC, *dBR, *dAO, *dBO;
    CK(hipMalloc(&A, bytes));
    CK(hipMalloc(&B, bytes));
    CK(hipMemset(A, 0, bytes));
    CK(hipMemset(B, 0, bytes));
    CK(hipMalloc(&dAC, VC * 8));
    CK(hipMalloc(&dBR, VR * 8));
    CK(hipMalloc(&dAO, VO * 8));
    CK(hipMalloc(&dBO, VO * 8));
    CK(hipMemcpy(dAC, tb.AC.data(), VC * 8, hipMemcpyHostToDevice));
    CK(hipMemcpy(dBR, tb.BR.data(), VR * 8, hipMemcpyHostToDevice));
    CK(hipMemcpy(dAO, tb.AO.data(), VO * 8, hipMemcpyHostToDevice));
    CK(hipMemcpy(dBO, tb.BO.data(), VO * 8, hipMemcpyHostToDevice));
    const char *mname[] = {"pattern", "no-barrier", "sequential"};
    // LDS reserved per block (unused) sets the blocks per CU: the tile
    // kernel's own reservation, and for the first workload 3 and 2 blocks
    // of 512 threads per CU.
    const uint32_t kern_lds = w.lds;
    std::vector<uint32_t> ldss{kern_lds};
    if (wi == 0 && !quick) {
      ldss.push_back(50000);
      ldss.push_back(72000);
    }
    for (uint32_t lds : ldss)
      for (int mode = 0; mode < 3; ++mode)
        for (int order = 0; order < 3; ++order) {
          if (mode == 2 && order > 0) continue;
          if ((lds != kern_lds || wi >= 3) && (mode == 1 || order > 0)) continue;
          auto launch = [&] {
#define HPTTX_PROBE(M, T)                                                                   \
  hipLaunchKernelGGL((k_pattern<M, T>), dim3(d.ntiles), dim3(T), lds, 0, A, B, dAC, dBR, dAO, \
                     dBO, d, order, 2.3f, 4.2f)
            if (w.nth == 512) {
              if (mode == 0) HPTTX_PROBE(0, 512);
              else if (mode == 1) HPTTX_PROBE(1, 512);
              else HPTTX_PROBE(2, 512);
            } else {
              if (mode == 0) HPTTX_PROBE(0, 256);
              else if (mode == 1) HPTTX_PROBE(1, 256);
              else HPTTX_PROBE(2, 256);
            }
#undef HPTTX_PROBE
            CK(hipGetLastError());
          };
          for (int r = 0; r < 3; ++r) launch();
          CK(hipDeviceSynchronize());
          const int reps = 20;
          CK(hipEventRecord(e0, 0));
          for (int r = 0; r < reps; ++r) launch();
          CK(hipEventRecord(e1, 0));
          CK(hipEventSynchronize(e1));
          float ms = 0;
          CK(hipEventElapsedTime(&ms, e0, e1));
          ms /= reps;
          std::printf("%-16s %-10s order %d  %u threads  lds %6u  %7.1f GB/s  (%.4f ms)\n",
                      w.name, mname[mode], order, w.nth, lds, 3.0 * bytes / (ms * 1e-3) / 1e9,
                      ms);
          std::fflush(stdout);
        }
    CK(hipFree(A));
    CK(hipFree(B));
    CK(hipFree(dAC));
    CK(hipFree(dBR));
    CK(hipFree(dAO));
    CK(hipFree(dBO));
  }
  return 0;
}
