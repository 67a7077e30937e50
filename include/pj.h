/*
 * pj.h -- C ABI of the MI355X-native radix-partitioned hash join (libpj.so).
 *
 * Drop-in boundary for the reference's partitioned-join path.  The reference
 * has no library API: its "entry point" is the set of host templates in
 * src/partitioned_join.cuh and src/skewed_data_handler.cuh called from
 * main.cu:759-893.  Each declaration below names the reference interface it
 * replaces.  All pointers named d_* or described as device pointers refer to
 * device (HBM) memory of the current HIP device; sizes are element counts
 * unless stated.  Functions never allocate device memory: scratch comes from
 * the caller's workspace.  Functions return a pj_status and never print or
 * abort; pj_last_error() returns a message for the calling thread's last
 * failure.
 *
 * Data layout
 *  - Tables are SoA like the reference (src/common.cuh:23-43): a key column of
 *    int32 or int64 (key_bytes 4 or 8) and an optional int32 payload column.
 *    A NULL payload means "row id" (payload[i] == i).  Inputs are read-only.
 *    Payload columns are int32 for both key widths.  The reference's 8-byte
 *    key build carries 8-byte payloads (value_type long long,
 *    src/common.cuh:26-27); here such columns are joined by row id (NULL
 *    payload) and fetched with pj_gather, which moves 4- or 8-byte columns.
 *  - Output pairs are {int32 build, int32 probe} with the int2 layout: the
 *    build and probe payloads of every matching (build row, probe row), i.e.
 *    (build_idx, probe_idx) when payloads are row ids (the cuDF gather-map
 *    contract).  Pair order is unspecified (the reference's is too:
 *    src/hash_join_kernels.cuh:185-200 claims output ranges with atomics).
 *  - Partitioned intermediate tuples ("pj tuples") are AoS: for 4-byte keys one
 *    8-byte word {int32 key, int32 payload}; for 8-byte keys 16 bytes
 *    {int64 key, int32 payload, int32 pad}.
 *  - Partition digits use raw low key bits like the reference: the final
 *    partition of a key is key & (2^(bits1+bits2) - 1); pass 1 uses bits
 *    [bits2, bits2+bits1), pass 2 bits [0, bits2)
 *    (src/partition_kernels.cuh:109,168,380).
 */
#ifndef PJ_H
#define PJ_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pj_status {
  PJ_OK = 0,
  PJ_ERR_INVALID_ARG = 1, /* bad pointer, size, key width or plan              */
  PJ_ERR_CAPACITY = 2,    /* output too small: *out_count holds the pair count */
  PJ_ERR_WORKSPACE = 3,   /* workspace smaller than pj_workspace_size()         */
  PJ_ERR_HIP = 4,         /* a HIP runtime call failed                          */
  PJ_ERR_UNSUPPORTED = 5  /* valid request outside what this build implements  */
} pj_status;

typedef struct pj_table {
  const void* keys;       /* device pointer: n keys of key_bytes each          */
  const int32_t* payload; /* device pointer or NULL (payload = row id)         */
  int64_t n;              /* rows, 0 <= n < 2^31                                */
  int32_t key_bytes;      /* 4 (int32 keys) or 8 (int64 keys)                   */
  int32_t reserved;
} pj_table;

typedef struct pj_pair {
  int32_t build_idx;
  int32_t probe_idx;
} pj_pair;

/* Join plan.  pj_plan_default() fills the plain histogram-driven plan; callers
 * may set fields explicitly (the reference's -n/-N/-R/-f/-u flags,
 * main.cu:100-209).  The plan never changes the result, only how it is made. */
typedef struct pj_plan {
  int32_t passes;            /* 1 or 2 radix partition passes                       */
  int32_t bits1;             /* radix bits of pass 1 (P1 = 2^bits1)                 */
  int32_t bits2;             /* radix bits of pass 2 (P2 = 2^bits2), 0 if passes==1 */
  int32_t unique_build_keys; /* 1: build keys are unique (reference -u): a probe
                                stops at its first match (hash_join_kernels.cuh:167) */
  double max_partition_size_factor; /* f: build chunk per workgroup is
                                       ceil(f*|R|/P) tuples (main.cu:867), clamped
                                       to what fits in LDS                        */
  int64_t probe_slice;       /* probe tuples per join work item; 0 = automatic      */
  int32_t source;            /* 0 = plain plan, 1 = learned planner, 2 = caller     */
  int32_t reserved[7];
} pj_plan;

/* Per-phase device times of the last pj_partitioned_join* call on this thread
 * when profiling is enabled (pj_profile_enable).  Phases match the columns of
 * the reference's CSV (scripts/bench_partitioned_join.sh:12). */
typedef struct pj_phase_times {
  float histogram_ms;   /* combined histogram of both tables (one key read)   */
  float offsets_ms;     /* column scans and partition offsets                 */
  float pass1_ms;       /* partition pass 1 (build + probe, one launch)       */
  float pass2_ms;       /* partition pass 2 (build + probe, one launch)       */
  float skew_ms;        /* join work list (block assignment)                  */
  float join_ms;        /* build + probe of all partitions (one launch)       */
  float compact_ms;     /* output compaction after the join                   */
  float total_ms;       /* first event to last event                          */
  int32_t valid;
  int32_t reserved;
} pj_phase_times;

/* ------------------------------------------------------------------------ */
/* Library and plans (host only; no device work)                             */
/* ------------------------------------------------------------------------ */
const char* pj_version(void);
/* Message of the calling thread's last error ("" if none). */
const char* pj_last_error(void);

/* Plain plan for a join of these tables (reads only n and key_bytes): the fewest radix bits whose
 * average build partition holds at most 2048 tuples (4-byte keys) or 1024 (8-byte keys), 1 pass
 * up to 2^10 partitions, else 2 passes (ceil / floor halves of the bits).
 * Replaces the reference's hand-chosen -n/-N (main.cu:239-254). */
pj_status pj_plan_default(const pj_table* build, const pj_table* probe, pj_plan* plan);

/* Bytes of workspace pj_partitioned_join needs for these tables and plan
 * (reads only n and key_bytes; plan NULL = plain plan). */
pj_status pj_workspace_size(const pj_table* build, const pj_table* probe, const pj_plan* plan,
                            size_t* bytes);

/* ------------------------------------------------------------------------ */
/* The join                                                                  */
/* ------------------------------------------------------------------------ */
/* Full radix-partitioned inner equi-join.  Replaces the reference's timed
 * sequence main.cu:759-893: combined_histogram_pass, partition passes one and
 * two, handle_skewed_data and shared_memory_hash_join.  plan NULL = plain plan.
 * Writes at most out_capacity pairs to out (device).  Synchronises `stream` at
 * the end and stores the total number of matching pairs in *out_count (host);
 * returns PJ_ERR_CAPACITY when it exceeds out_capacity (the contents of out
 * are then unspecified).  Unlike the reference it
 * needs no CPU-derived output size (main.cu:687-708) and never overwrites its
 * inputs (main.cu:833-852). */
pj_status pj_partitioned_join(const pj_table* build, const pj_table* probe, const pj_plan* plan,
                              pj_pair* out, int64_t out_capacity, int64_t* out_count,
                              void* workspace, size_t workspace_bytes, hipStream_t stream);

/* Stream-ordered form: no host synchronisation, no host reads; the pair count
 * is written to d_out_count (device int64).  Graph-capturable. */
pj_status pj_partitioned_join_async(const pj_table* build, const pj_table* probe, const pj_plan* plan,
                                    pj_pair* out, int64_t out_capacity, int64_t* d_out_count,
                                    void* workspace, size_t workspace_bytes, hipStream_t stream);

/* Allocation-free form of pj_partitioned_join_async.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The workspace is a static device array of 192 MiB inside libpj.so, one per
 * GPU, held until the process ends; joins whose pj_workspace_size exceeds it
 * return PJ_ERR_WORKSPACE.  Nothing is allocated or freed.  Every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU (other
 * streams, other threads) race: serialise them. */
pj_status pj_partitioned_join_noalloc(const pj_table* build, const pj_table* probe, const pj_plan* plan,
                                      pj_pair* out, int64_t out_capacity, int64_t* d_out_count,
                                      hipStream_t stream);

/* Projection to the reference's output form {x = build key, y = probe payload}
 * (src/hash_join_kernels.cuh:155-164): out_xy[2i] = key of build row
 * pairs[i].build_idx, out_xy[2i+1] = pairs[i].probe_idx.  Requires pairs made
 * with row-id build payloads.  out_xy has key_bytes-wide elements. */
pj_status pj_project_reference(const pj_table* build, const pj_pair* pairs, int64_t n, void* out_xy,
                               hipStream_t stream);

/* Gather map: out_build[i] = build_col[pairs[i].build_idx] and out_probe[i] =
 * probe_col[pairs[i].probe_idx] for i < n, for columns of 4- or 8-byte
 * elements; either column may be NULL (not gathered).  Requires pairs made
 * with row-id payloads.  The reference returns no indices to gather with; this
 * is the consumer of cuDF's inner_join gather maps
 * (cudf:include/cudf/join.hpp:95-100, cudf:src/join/join.cu:33-64). */
pj_status pj_gather(const pj_pair* pairs, int64_t n, const void* build_col, int32_t build_bytes,
                    const void* probe_col, int32_t probe_bytes, void* out_build, void* out_probe,
                    hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* Phase-level entry points (each mirrors one reference host function)        */
/* ------------------------------------------------------------------------ */
/* Combined histogram: d_fine[q] = #{i : key_i & (2^total_bits - 1) == q} and
 * d_pass1[p] = sum of d_fine over q >> bits2 == p.  Replaces
 * combined_histogram_pass (src/partitioned_join.cuh:22-48: big_histogram +
 * retrieve_histogram_of_pass_one).  total_bits <= 16.  Needs
 * pj_histogram_workspace_size() bytes of workspace. */
pj_status pj_combined_histogram(const pj_table* table, int32_t total_bits, int32_t bits2, uint32_t* d_fine,
                                uint32_t* d_pass1, void* workspace, size_t workspace_bytes,
                                hipStream_t stream);
size_t pj_histogram_workspace_size(int32_t total_bits);

/* Digit histogram d_hist[d] = #{i : (key_i >> shift) & (2^bits-1) == d}.
 * Replaces histogram_pass_one (src/partition_kernels.cuh:140-194) and, with
 * shift = 0, the per-chunk histogram of histogram_pass_two (:328-406). */
pj_status pj_histogram(const pj_table* table, int32_t shift, int32_t bits, uint32_t* d_hist,
                       void* workspace, size_t workspace_bytes, hipStream_t stream);

/* Exclusive / inclusive prefix sums of n uint32 (results uint32, wrap mod 2^32).
 * Replace cub::DeviceScan::ExclusiveSum / InclusiveSum
 * (src/partitioned_join.cuh:86,152; src/skewed_data_handler.cuh:79-90). */
pj_status pj_exclusive_scan(const uint32_t* d_in, uint32_t* d_out, int64_t n, void* workspace,
                            size_t workspace_bytes, hipStream_t stream);
pj_status pj_inclusive_scan(const uint32_t* d_in, uint32_t* d_out, int64_t n, void* workspace,
                            size_t workspace_bytes, hipStream_t stream);
size_t pj_scan_workspace_size(int64_t n);

/* One radix partition pass of a whole table into pj tuples: digit
 * (key >> shift) & (2^bits - 1), partitions laid out in digit order; d_offsets
 * (2^bits + 1 entries) receives partition starts and n.  Replaces
 * partition_input_tables_pass_one (src/partitioned_join.cuh:50-107). */
pj_status pj_partition_pass(const pj_table* table, int32_t shift, int32_t bits, void* d_tuples_out,
                            uint32_t* d_offsets, void* workspace, size_t workspace_bytes,
                            hipStream_t stream);
size_t pj_partition_workspace_size(const pj_table* table, int32_t bits);

/* Second pass: d_tuples_in holds n pj tuples grouped into 2^bits1 partitions
 * at d_offsets1 (2^bits1 + 1 entries); each partition p1 is split by digit
 * key & (2^bits2 - 1) into d_tuples_out, final partition p1*2^bits2 + p2 starting
 * at d_offsets2[p1*2^bits2 + p2] (2^(bits1+bits2) + 1 entries).  Replaces
 * partition_input_tables_pass_two (src/partitioned_join.cuh:109-175). */
pj_status pj_partition_pass2(const void* d_tuples_in, int64_t n, int32_t key_bytes, int32_t bits1,
                             int32_t bits2, const uint32_t* d_offsets1, void* d_tuples_out,
                             uint32_t* d_offsets2, void* workspace, size_t workspace_bytes,
                             hipStream_t stream);
size_t pj_partition2_workspace_size(int64_t n, int32_t bits1, int32_t bits2);

/* Block assignment of handle_skewed_data (src/skewed_data_handler.cuh:58-91)
 * without its host round trip: for partitions with exclusive offsets
 * d_offsets[0..P) of a table of `total` rows, blocks[p] = max(1, ceil(size/base));
 * d_blocks_scanned[p] = inclusive scan of blocks, d_partition_index[b] = p
 * owning block b (capacity max_blocks entries), *d_total_blocks = sum. */
pj_status pj_assign_blocks(const uint32_t* d_offsets, int64_t total, int32_t num_partitions, int32_t base,
                           uint32_t* d_blocks_scanned, uint32_t* d_partition_index, int64_t max_blocks,
                           uint32_t* d_total_blocks, void* workspace, size_t workspace_bytes, hipStream_t stream);
size_t pj_assign_blocks_workspace_size(int32_t num_partitions);

/* Join of already partitioned tables: build and probe pj tuples grouped into
 * P = 2^total_bits partitions at d_build_offsets / d_probe_offsets (P + 1
 * entries each).  Builds an LDS open-addressing table (MurmurHash3 seed 0,
 * int_fastdiv modulo, linear probing, occupancy <= 0.5) per (partition, build
 * chunk of <= base tuples) and probes it with probe slices of <= probe_slice
 * tuples; heavy partitions are split on both sides.  Replaces
 * shared_memory_hash_join + hash_join (src/partitioned_join.cuh:177-251,
 * src/hash_join_kernels.cuh:23-201).  Pair count goes to d_out_count. */
pj_status pj_join_partitions(const void* d_build_tuples, const uint32_t* d_build_offsets,
                             const void* d_probe_tuples, const uint32_t* d_probe_offsets, int32_t key_bytes,
                             int32_t total_bits, int32_t base, int64_t probe_slice, int32_t unique_build_keys,
                             pj_pair* out, int64_t out_capacity, int64_t* d_out_count, void* workspace,
                             size_t workspace_bytes, hipStream_t stream);
size_t pj_join_workspace_size(int32_t total_bits);

/* ------------------------------------------------------------------------ */
/* Device primitives exposed for parity tests                                */
/* ------------------------------------------------------------------------ */
/* d_out[i] = MurmurHash3_x86_32(seed 0) of key i (src/hash_functions.cuh:33-125). */
pj_status pj_hash_keys(const void* d_keys, int64_t n, int32_t key_bytes, uint32_t* d_out, hipStream_t stream);
/* d_q[i] = n_i / d and d_r[i] = n_i % d with int_fastdiv (src/int_fastdiv.h:102-123). */
pj_status pj_fastdiv(const int32_t* d_n, int64_t count, int32_t divisor, int32_t* d_q, int32_t* d_r,
                     hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* Seeded synthetic generators (device).  Own definitions: SplitMix64
 * counters and a keyed bijection; semantics follow the reference generator
 * (src/generate_input_tables.cuh:116-141: unique lottery, selectivity) and
 * Zipf sampler (src/generate_zipf_input.cuh:50-127).                        */
/* ------------------------------------------------------------------------ */
/* keys[i] = perm_domain(i), i < n: n distinct keys drawn from [0, domain). */
pj_status pj_gen_build_unique_i32(int32_t* d_keys, int64_t n, int64_t domain, uint64_t perm_seed,
                                  hipStream_t stream);
/* keys[i] = perm_domain(U(0, m)): keys with repeats. */
pj_status pj_gen_build_dup_i32(int32_t* d_keys, int64_t n, int64_t m, int64_t domain, uint64_t perm_seed,
                               uint64_t seed, hipStream_t stream);
/* With probability sel a build key perm_domain(U(0, n_build)); otherwise
 * perm_domain(n_build + U(0, domain - n_build)) if nonmatch_range == 0, else
 * nonmatch_base + U(0, nonmatch_range). */
pj_status pj_gen_probe_i32(int32_t* d_keys, int64_t n, int64_t domain, int64_t n_build, uint64_t perm_seed,
                           uint64_t seed, double sel, int64_t nonmatch_base, int64_t nonmatch_range,
                           hipStream_t stream);
/* Zipf CDF table on the host: lut[i-1] = H(i, theta) / H(n, theta). */
pj_status pj_zipf_lut(double* h_lut, int64_t n, double theta);
/* Probe keys perm_n(rank), rank ~ Zipf over [0, n) by binary search in the
 * device copy d_lut of pj_zipf_lut. */
pj_status pj_gen_probe_zipf_i32(int32_t* d_keys, int64_t count, const double* d_lut, int64_t n,
                                uint64_t rank_perm_seed, uint64_t seed, hipStream_t stream);
pj_status pj_gen_build_i64(int64_t* d_keys, int64_t n, uint64_t key_seed, hipStream_t stream);
pj_status pj_gen_probe_i64(int64_t* d_keys, int64_t n, int64_t n_build, uint64_t key_seed, uint64_t seed,
                           double sel, hipStream_t stream);
pj_status pj_gen_probe_zipf_i64(int64_t* d_keys, int64_t count, const double* d_lut, int64_t n_build,
                                uint64_t key_seed, uint64_t rank_perm_seed, uint64_t seed, hipStream_t stream);
/* d_pay[i] = i */
pj_status pj_gen_row_ids(int32_t* d_pay, int64_t n, hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* Profiling (the reference's per-phase cudaEvent timing, main.cu:769-896)    */
/* ------------------------------------------------------------------------ */
/* on != 0: record hipEvents around each phase of later joins on this thread. */
pj_status pj_profile_enable(int32_t on);
/* Synchronises on the last recorded events and returns the phase times. */
pj_status pj_profile_read(pj_phase_times* times);
/* The phase boundaries later profiled joins on this thread record: bit i
 * (0..7) is the event before phase i (bit 7: after the last phase); default
 * 0xFF.  Each event costs the GPU a few microseconds, so a timing loop may
 * record only the two around the phase it times.  Phases without both of
 * their events read 0, as does total_ms without events 0 and 7. */
pj_status pj_profile_select(uint32_t marks);
/* The phase times of the last min(n, 128, joins profiled since
 * pj_profile_enable) profiled joins on this thread, oldest first, into
 * times[0..*got), after one wait for the newest: joins queued back to back
 * are timed without a host wait between them.  Joins with an empty table
 * record no times. */
pj_status pj_profile_read_recent(pj_phase_times* times, int32_t n, int32_t* got);

/* ------------------------------------------------------------------------ */
/* Learned planner.  No reference counterpart: the reference's plan is fixed  */
/* by its -n/-N/-f flags (main.cu:100-209).  The planner only chooses a plan; */
/* joins made with any plan give the same pairs.                             */
/*  - features: 1030 floats = a 4096-key probe sample's 1024-bin hash        */
/*    histogram (bins sorted by count, as fractions) + log2|R|/32, log2|S|/32,*/
/*    log2(|S|/|R|)/8, key_bytes/8, top sampled key frequency, sample        */
/*    collision probability;                                                */
/*  - candidates: the 64 most frequent sampled keys (count 0 = no candidate);*/
/*  - MLP 1030 -> 256 -> 256 -> 81 logits (ReLU): bits of pass 1 (6..12),    */
/*    bits of pass 2 (6..12), pass count (1..3), 64 heavy-key flags;        */
/*  - weights: pj_planner_param_count() floats, W1 [1030x256], b1, W2         */
/*    [256x256], b2, W3 [256x81], b3, row-major.                           */
/* ------------------------------------------------------------------------ */
size_t pj_planner_param_count(void);
/* Largest build chunk of one join work item (LDS table capacity / 2 - 1); a
 * plan with ceil(f * |R| / P) above it splits every partition's build side. */
int32_t pj_join_chunk_capacity(int32_t key_bytes);
size_t pj_planner_scratch_bytes(int32_t batch);
pj_status pj_planner_sketch(const pj_table* build, const pj_table* probe, float* d_features, uint64_t* d_cand_keys,
                            uint32_t* d_cand_sample_counts, hipStream_t stream);
/* Exact occurrences of the candidates in the probe column (d_counts[64]). */
pj_status pj_planner_candidate_counts(const pj_table* probe, const uint64_t* d_cand_keys,
                                      const uint32_t* d_cand_sample_counts, uint64_t* d_counts, hipStream_t stream);
/* logits[batch][81] for features[batch][1030]. */
pj_status pj_planner_forward(const float* d_weights, const float* d_features, int32_t batch, float* d_logits,
                             void* d_scratch, size_t scratch_bytes, hipStream_t stream);
/* One Adam step (beta 0.9/0.999) on the mean loss of a batch: softmax cross-
 * entropy on the three plan heads (labels[b] = {bits1 - 6, bits2 - 6 or -1,
 * passes - 1}) plus binary cross-entropy on the heavy flags (heavy[b][j] = 1,
 * 0, or -1 where there is no candidate); *d_loss receives the mean loss. */
pj_status pj_planner_train_step(float* d_weights, float* d_adam_m, float* d_adam_v, int32_t step, float lr,
                                const float* d_features, const int32_t* d_labels, const float* d_heavy,
                                int32_t batch, float* d_loss, void* d_scratch, size_t scratch_bytes,
                                hipStream_t stream);
/* Host only: the plan from one sample's logits.  A key is heavy with at least
 * T = |S| / 1024 probe rows.  The predicted heavy set (flag logit > 0)
 * disagrees with the exact counts if a candidate predicted heavy has fewer
 * than T / 2 rows or one not predicted heavy has 2T or more.  On any
 * disagreement, or if the plan's partitions are larger than the plain plan's
 * aim and its build chunks would not fit the LDS table, the plain plan is
 * returned and *fallback = 1. */
pj_status pj_plan_decode(const float* logits, const uint32_t* cand_sample_counts, const uint64_t* cand_exact_counts,
                         const pj_table* build, const pj_table* probe, pj_plan* plan, int32_t* fallback);
size_t pj_plan_learned_workspace_size(void);
/* Sketch, forward pass, exact candidate counts, then pj_plan_decode.  When the
 * predicted plan is the plain plan itself the heavy-set check has nothing to
 * choose between, so the counts are skipped (*fallback = 0, plan = that plan).
 * Synchronises `stream`. */
pj_status pj_plan_learned(const pj_table* build, const pj_table* probe, const float* d_weights, pj_plan* plan,
                          int32_t* fallback, void* workspace, size_t workspace_bytes, hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* Batched joins.  No reference counterpart (the reference runs one join per  */
/* process).  Many independent joins are captured once into one HIP graph:    */
/* the joins are spread over `chains` parallel chains (longest first by      */
/* |R| + |S|), each chain runs its joins' pipelines one after another on its */
/* own region of the workspace, and one pj_batch_launch replays the whole    */
/* graph.  Join j writes at most out_caps[j] pairs at out + out_offsets[j]    */
/* and its pair count to d_counts[j] (device), as pj_partitioned_join_async  */
/* does.  The tables, output and workspace must stay allocated and unchanged */
/* in address while the batch exists; their contents may change between      */
/* launches.                                                                 */
/* ------------------------------------------------------------------------ */
typedef struct pj_batch pj_batch;
/* plans: NULL (plain plan for every join) or one plan per join. */
pj_status pj_batch_workspace_size(const pj_table* builds, const pj_table* probes, const pj_plan* plans,
                                  int32_t count, int32_t chains, size_t* bytes);
pj_status pj_batch_create(const pj_table* builds, const pj_table* probes, const pj_plan* plans, int32_t count,
                          int32_t chains, pj_pair* out, const int64_t* out_offsets, const int64_t* out_caps,
                          int64_t* d_counts, void* workspace, size_t workspace_bytes, pj_batch** batch);
pj_status pj_batch_launch(pj_batch* batch, hipStream_t stream);
pj_status pj_batch_destroy(pj_batch* batch);

#ifdef __cplusplus
}
#endif
#endif /* PJ_H */
