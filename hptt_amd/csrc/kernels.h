// kernels.h -- launch entry for one planned transposition.
#pragma once

#include <hip/hip_runtime.h>

#include "planner.h"

namespace hpttx {

// Coefficients as the user gave them (real: ai = bi = 0).
struct CoefD {
  double ar, ai, br, bi;
};

// Enqueue the kernel for `lp` on `st`.  `lp` must be finalized with vector
// widths legal for A and B (finalize_vectors).
hipError_t launch_plan(int dtype, const LaunchPlan &lp, const void *A, void *B,
                       const CoefD &c, bool update, bool cc, hipStream_t st);

uint32_t vec_cap(int dtype);

}  // namespace hpttx
