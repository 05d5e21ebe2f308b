// tune.h -- tile/schedule autotuner with an in-process and on-disk cache.
//
// Replaces the reference's timed tuning over loop orders and thread splits
// (inc/hptt/plan/plan_trans.tcc:33-67, candidates from
// plan_trans_util.tcc:1183-1266) keeping its tuning_timeout meaning
// (inc/hptt/hptt.h:46-57): 0 = model pick, > 0 = time-boxed, < 0 = every
// candidate.  Unlike the reference, which times candidates on the caller's
// buffers with alpha=0, beta=1 (plan_trans.tcc:42-44), tuning runs on scratch
// device buffers and never touches the caller's B.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "kernels.h"
#include "planner.h"

namespace hpttx {

// Cache key: the current device's architecture, dtype, update and
// complex-coefficient flags and the fused problem (extents and strides).
std::string tune_key(const Problem &p, bool update, bool cc);

// Look the problem up in the cache (memory, then the file named by the
// HPTTX_TUNE_CACHE environment variable).
bool tune_lookup(const Problem &p, bool update, bool cc, TuneParams &out);

// Replace a cached entry that no longer plans (a stale or foreign record)
// with the model's pick, in memory and in the file.
void tune_forget(const std::string &key);

// Time candidate tile shapes and keep the fastest; records it in the cache.
// The candidates run on scratch copies of the caller's A and B (src_a /
// src_b, read only; zero-filled scratch when null), never on the caller's B.
int autotune(const Problem &p, const CoefD &coef, bool update, bool cc,
             double timeout_s, hipStream_t stream, TuneParams &out,
             std::string &err, const void *src_a = nullptr, const void *src_b = nullptr);

// The candidate list the tuner walks (model default first).
std::vector<TuneParams> tune_candidates(const Problem &p);

// The same cache under an explicit key (batch classes are tuned in batch
// context and cached under their own keys).
bool tune_lookup_key(const std::string &key, TuneParams &out);
void tune_store_key(const std::string &key, const TuneParams &t);

// Two launch plans with the same launch shape (tiles or rows, streaming).
bool same_launch(const LaunchPlan &a, const LaunchPlan &b);

}  // namespace hpttx
