// pj_planner.h -- learned join planner: shapes and launchers of pj_planner.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace pj {

constexpr int kPlanSample = 4096;  // probe keys sampled by the sketch
constexpr int kPlanBins = 1024;    // hash bins of the sample
constexpr int kPlanCand = 64;      // heavy-key candidates (most frequent sampled keys)
constexpr int kPlanFeat = kPlanBins + 6;
constexpr int kPlanHidden = 256;
constexpr int kPlanBitsCls = 7;    // radix bits 6..12 of one pass
constexpr int kPlanPassCls = 3;    // 1..3 passes
constexpr int kPlanOut = 2 * kPlanBitsCls + kPlanPassCls + kPlanCand;  // 81 logits

// Parameter block: W1 [F x H], b1 [H], W2 [H x H], b2 [H], W3 [H x O], b3 [O].
constexpr size_t kPlanW1 = 0;
constexpr size_t kPlanB1 = kPlanW1 + size_t(kPlanFeat) * kPlanHidden;
constexpr size_t kPlanW2 = kPlanB1 + kPlanHidden;
constexpr size_t kPlanB2 = kPlanW2 + size_t(kPlanHidden) * kPlanHidden;
constexpr size_t kPlanW3 = kPlanB2 + kPlanHidden;
constexpr size_t kPlanB3 = kPlanW3 + size_t(kPlanHidden) * kPlanOut;
constexpr size_t kPlanParams = kPlanB3 + kPlanOut;

hipError_t planner_sketch(const void* keys, int64_t n, int key_bytes, int64_t n_build, float* feat,
                          uint64_t* cand_keys, uint32_t* cand_counts, hipStream_t s);
hipError_t planner_count_candidates(const void* keys, int64_t n, int key_bytes, const uint64_t* cand_keys,
                                    const uint32_t* cand_sample_counts, unsigned long long* counts, hipStream_t s);
size_t planner_scratch_floats(int batch);
hipError_t planner_forward(const float* w, const float* x, int batch, float* out, float* scratch, hipStream_t s);
hipError_t planner_train_step(float* w, float* m, float* v, int step, float lr, const float* x, const int32_t* labels,
                              const float* heavy, int batch, float* loss, float* scratch, hipStream_t s);

}  // namespace pj
