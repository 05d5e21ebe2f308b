/*
 * hptt_oracle.c -- CPU restatement of HPTT's out-of-place transposition.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke entry
 * point and the benchmark's CPU-baseline leg may load it; the product library
 * (hptt_amd/) never links, loads or calls it.
 *
 * Operation (reference README.md:6-12, inc/hptt/hptt.h:18-22):
 *
 *     B[j] = alpha * A[i]            if !UPDATE
 *     B[j] = beta * B[j] + alpha*A[i] if UPDATE,   j_d = i_{perm[d]}
 *
 * Followed reference code, clause by clause:
 *   - column-major strides from the outer sizes: stride_0 = 1,
 *     stride_{d+1} = outer_d * stride_d          (inc/hptt/tensor.tcc:186-196)
 *   - output sizes out_size[d] = in_size[perm[d]]
 *                              (cmake/pylib/gen_util/hptt_trans_gen.py:196-197)
 *   - index map loop_perm_idx_[d] = &loop_idx_[perm[d]]
 *                              (inc/hptt/operations/operation_trans.tcc:47-51)
 *   - loop structure: one parallel loop over every non-leading output index,
 *     inner loop over the output's stride-1 dim with a strided read of A
 *                              (test/inc/hptt/test_util.tcc:143-185, RefTrans)
 *   - UPDATE iff fabs(beta) >= 1e-16 (double compare)
 *                              (inc/hptt/impl/hptt_trans_impl.h:75-78)
 *   - arithmetic order: UPDATE r = fl(fl(beta*y) + fl(alpha*x)), two rounded
 *     products and one rounded add, no fused multiply-add; !UPDATE r = fl(alpha*x)
 *     and B is never read (src/arch/avx2/kernel_trans_avx2.cc:67-108,
 *     inc/hptt/arch/common/common_impl.h:18-33).  Built with -ffp-contract=off.
 *   - complex dtypes with real coefficients: alpha/beta applied to the real and
 *     imaginary parts independently (inc/hptt/types.h:29-54).
 *   - argument validation, nullptr on failure (src/hptt.cc:25-56), with two
 *     deliberate differences: the outer-size check on B compares
 *     out_outer[d] >= in_size[perm[d]] (the reference indexes out_outer by
 *     perm[i] and compares against in_size[i], which is only right for
 *     involutive perms), and ranks outside [2,6] are rejected instead of
 *     producing a plan that silently does nothing.
 *
 * Extension not expressible in the reference (its coefficients are real):
 * complex alpha/beta for complex dtypes with the fixed order
 *     ax = (fl(fl(ar*xr) - fl(ai*xi)), fl(fl(ar*xi) + fl(ai*xr)))
 *     by = (fl(fl(br*yr) - fl(bi*yi)), fl(fl(br*yi) + fl(bi*yr)))
 *     r  = (fl(by.re + ax.re), fl(by.im + ax.im))
 * and UPDATE iff fabs(br) >= 1e-16 or fabs(bi) >= 1e-16.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_MAX_RANK 6

/* Planted mistakes, for the tests that show the checks have teeth.  The
 * shipped oracle is built with ORACLE_MUTANT=0; the Makefile's `mutants`
 * target builds liboracle_mutant<k>.so:
 *   1: inverse permutation in the index map   2: fused multiply-add
 *   3: beta*B term dropped                    4: UPDATE threshold ignored
 *   5: sign of the complex cross term flipped */
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

enum { ORACLE_OK = 0, ORACLE_EINVAL = 1, ORACLE_EUNSUPPORTED = 4 };

/* Mirrors src/hptt.cc:25-56 (see header for the two deliberate differences). */
static int oracle_validate(const void *A, const void *B, uint32_t rank,
                           const uint32_t *in_size, const uint32_t *perm,
                           const uint32_t *in_outer, const uint32_t *out_outer) {
  if (A == NULL || B == NULL) return ORACLE_EINVAL;
  if (rank <= 1) return ORACLE_EINVAL;
  if (rank > ORACLE_MAX_RANK) return ORACLE_EUNSUPPORTED;
  if (in_size == NULL || perm == NULL) return ORACLE_EINVAL;
  int seen[ORACLE_MAX_RANK] = {0};
  for (uint32_t d = 0; d < rank; ++d) {
    if (perm[d] >= rank || seen[perm[d]]) return ORACLE_EINVAL;
    seen[perm[d]] = 1;
  }
  for (uint32_t d = 0; d < rank; ++d) {
    if (in_size[d] == 0) return ORACLE_EINVAL;
    if (in_outer && in_outer[d] < in_size[d]) return ORACLE_EINVAL;
    if (out_outer && out_outer[d] < in_size[perm[d]]) return ORACLE_EINVAL;
  }
  return ORACLE_OK;
}

/* Column-major strides over the outer sizes (tensor.tcc:186-196). */
static void oracle_strides(uint32_t rank, const uint32_t *size,
                           const uint32_t *outer, uint64_t *stride) {
  stride[0] = 1;
  for (uint32_t d = 0; d + 1 < rank; ++d)
    stride[d + 1] = stride[d] * (uint64_t)(outer ? outer[d] : size[d]);
}

int oracle_update_real(double beta) { return fabs(beta) >= 1e-16; }

int oracle_update_cplx(double br, double bi) {
  return fabs(br) >= 1e-16 || fabs(bi) >= 1e-16;
}

/*
 * One body per real base type R.  NC = 1 (real) or 2 (complex, interleaved
 * re/im).  CC = 1 selects complex coefficients (NC must be 2).
 */
#define ORACLE_BODY(R)                                                         \
  int rc = oracle_validate(A, B, rank, in_size, perm, in_outer, out_outer);    \
  if (rc != ORACLE_OK) return rc;                                              \
  uint32_t pm[ORACLE_MAX_RANK];                                                \
  for (uint32_t d = 0; d < rank; ++d) pm[d] = perm[d];                         \
  if (ORACLE_MUTANT == 1)                                                      \
    for (uint32_t d = 0; d < rank; ++d) pm[perm[d]] = d;                       \
  perm = pm;                                                                   \
  uint32_t out_size[ORACLE_MAX_RANK];                                          \
  for (uint32_t d = 0; d < rank; ++d) out_size[d] = in_size[perm[d]];          \
  uint64_t sa[ORACLE_MAX_RANK], sb[ORACLE_MAX_RANK];                           \
  oracle_strides(rank, in_size, in_outer, sa);                                 \
  oracle_strides(rank, out_size, out_outer, sb);                               \
  const uint64_t n_inner = out_size[0];                                        \
  const uint64_t sa_inner = sa[perm[0]];                                       \
  uint64_t n_outer = 1;                                                        \
  for (uint32_t d = 1; d < rank; ++d) n_outer *= out_size[d];                  \
  const R ar = (R)alpha_re, ai = (R)alpha_im, br = (R)beta_re,                 \
          bi = (R)beta_im;                                                     \
  int update = cplx_coef ? oracle_update_cplx(br, bi)                          \
                         : oracle_update_real(br);                             \
  if (ORACLE_MUTANT == 4) update = (br != 0 || bi != 0);                       \
  if (nthreads <= 0) nthreads = 1;                                             \
  _Pragma("omp parallel for schedule(static) num_threads(nthreads)")          \
  for (uint64_t idx = 0; idx < n_outer; ++idx) {                               \
    uint64_t off_a = 0, off_b = 0, t = idx;                                    \
    for (uint32_t d = 1; d < rank; ++d) {                                      \
      uint64_t j = t % out_size[d];                                            \
      t /= out_size[d];                                                        \
      off_a += j * sa[perm[d]];                                                \
      off_b += j * sb[d];                                                      \
    }                                                                          \
    for (uint64_t k = 0; k < n_inner; ++k) {                                   \
      const R *x = A + (off_a + k * sa_inner) * (uint64_t)nc;                  \
      R *y = B + (off_b + k) * (uint64_t)nc;                                   \
      if (!cplx_coef) {                                                        \
        for (int c = 0; c < nc; ++c) {                                         \
          R ax = ar * x[c];                                                    \
          if (update) {                                                        \
            R by = br * y[c];                                                  \
            if (ORACLE_MUTANT == 2) y[c] = (R)fma(br, y[c], ax);              \
            else if (ORACLE_MUTANT == 3) y[c] = ax;                            \
            else y[c] = by + ax;                                               \
          } else {                                                             \
            y[c] = ax;                                                         \
          }                                                                    \
        }                                                                      \
      } else {                                                                 \
        R p0 = ar * x[0], p1 = ai * x[1], p2 = ar * x[1], p3 = ai * x[0];      \
        R axr = (ORACLE_MUTANT == 5) ? p0 + p1 : p0 - p1, axi = p2 + p3;       \
        if (update) {                                                          \
          R q0 = br * y[0], q1 = bi * y[1], q2 = br * y[1], q3 = bi * y[0];    \
          R byr = q0 - q1, byi = q2 + q3;                                      \
          y[0] = byr + axr;                                                    \
          y[1] = byi + axi;                                                    \
        } else {                                                               \
          y[0] = axr;                                                          \
          y[1] = axi;                                                          \
        }                                                                      \
      }                                                                        \
    }                                                                          \
  }                                                                            \
  return ORACLE_OK;

static int oracle_f32(const float *A, float *B, int nc, int cplx_coef,
                      uint32_t rank, const uint32_t *in_size,
                      const uint32_t *perm, double alpha_re, double alpha_im,
                      double beta_re, double beta_im, const uint32_t *in_outer,
                      const uint32_t *out_outer, int nthreads) {
  ORACLE_BODY(float)
}

static int oracle_f64(const double *A, double *B, int nc, int cplx_coef,
                      uint32_t rank, const uint32_t *in_size,
                      const uint32_t *perm, double alpha_re, double alpha_im,
                      double beta_re, double beta_im, const uint32_t *in_outer,
                      const uint32_t *out_outer, int nthreads) {
  ORACLE_BODY(double)
}

/*
 * Public entry points.  dtype: 0 = float, 1 = double, 2 = complex<float>,
 * 3 = complex<double>.  cplx_coef != 0 selects the complex-coefficient form
 * (dtype 2/3 only).  in_outer / out_outer may be NULL (= unpadded).
 */
int oracle_transpose(int dtype, const void *A, void *B, uint32_t rank,
                     const uint32_t *in_size, const uint32_t *perm,
                     double alpha_re, double alpha_im, double beta_re,
                     double beta_im, int cplx_coef, const uint32_t *in_outer,
                     const uint32_t *out_outer, int nthreads) {
  switch (dtype) {
    case 0:
      if (cplx_coef) return ORACLE_EINVAL;
      return oracle_f32((const float *)A, (float *)B, 1, 0, rank, in_size, perm,
                        alpha_re, 0.0, beta_re, 0.0, in_outer, out_outer,
                        nthreads);
    case 1:
      if (cplx_coef) return ORACLE_EINVAL;
      return oracle_f64((const double *)A, (double *)B, 1, 0, rank, in_size,
                        perm, alpha_re, 0.0, beta_re, 0.0, in_outer, out_outer,
                        nthreads);
    case 2:
      return oracle_f32((const float *)A, (float *)B, 2, cplx_coef != 0, rank,
                        in_size, perm, alpha_re, alpha_im, beta_re, beta_im,
                        in_outer, out_outer, nthreads);
    case 3:
      return oracle_f64((const double *)A, (double *)B, 2, cplx_coef != 0, rank,
                        in_size, perm, alpha_re, alpha_im, beta_re, beta_im,
                        in_outer, out_outer, nthreads);
    default:
      return ORACLE_EINVAL;
  }
}

/* Deterministic input generator shared by tests and the benchmark: a
 * splitmix64 counter stream mapped to uniform[lo, hi).  Element k of the
 * scalar stream depends only on (seed, k), so any thread split gives the same
 * data (the reference's generator draws from one shared mt19937 inside an
 * OpenMP loop, test/inc/hptt/test_util.tcc:31-41, which is neither
 * reproducible nor race-free). */
static inline uint64_t oracle_splitmix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

void oracle_fill_uniform_f32(float *dst, uint64_t n, uint64_t seed, double lo,
                             double hi, int nthreads) {
  if (nthreads <= 0) nthreads = 1;
  const uint64_t key = oracle_splitmix(seed);
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (uint64_t k = 0; k < n; ++k) {
    uint64_t r = oracle_splitmix(key ^ (k * 0xD1B54A32D192ED03ull));
    double u = (double)(r >> 40) * (1.0 / 16777216.0); /* 24 bits */
    dst[k] = (float)(lo + (hi - lo) * u);
  }
}

void oracle_fill_uniform_f64(double *dst, uint64_t n, uint64_t seed, double lo,
                             double hi, int nthreads) {
  if (nthreads <= 0) nthreads = 1;
  const uint64_t key = oracle_splitmix(seed);
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (uint64_t k = 0; k < n; ++k) {
    uint64_t r = oracle_splitmix(key ^ (k * 0xD1B54A32D192ED03ull));
    double u = (double)(r >> 11) * (1.0 / 9007199254740992.0); /* 53 bits */
    dst[k] = lo + (hi - lo) * u;
  }
}

/* Bitwise comparison helper: returns the index of the first differing byte
 * block of `elem_bytes`, or -1 when equal. */
int64_t oracle_first_mismatch(const void *x, const void *y, uint64_t n,
                              uint32_t elem_bytes) {
  const unsigned char *a = (const unsigned char *)x;
  const unsigned char *b = (const unsigned char *)y;
  for (uint64_t k = 0; k < n; ++k)
    if (memcmp(a + k * elem_bytes, b + k * elem_bytes, elem_bytes) != 0)
      return (int64_t)k;
  return -1;
}
