/*
 * pj_oracle.h -- CPU restatement of the reference radix-partitioned join path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library (libpj.so) links,
 * loads or calls this code.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).  Each function is pinned
 * by an independent check (closed form, brute force, published verification
 * value or a known-answer test); the pins live in tests/test_oracle_*.py.
 */
#ifndef PJ_ORACLE_H
#define PJ_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- MurmurHash3_x86_32 (src/hash_functions.cuh:33-125) ---------------- */
/* General byte-string form (Appleby, SMHasher); pinned by the SMHasher
 * verification value 0xB0F57EE3. */
uint32_t pjo_murmur3_32(const void* data, int len, uint32_t seed);
/* Seed-0 hash of one 4- or 8-byte key, as src/hash_functions.cuh:80-119 applies
 * it (little-endian key bytes, len mixed into the finaliser). */
uint32_t pjo_murmur3_key4(uint32_t key);
uint32_t pjo_murmur3_key8(uint64_t key);

/* ---- int_fastdiv (src/int_fastdiv.h:21-133) ------------------------------ */
/* Signed 32-bit division by a runtime-invariant divisor with Hacker's Delight
 * (2nd ed., ch. 10) magic numbers.  Pinned against C '/' and '%'. */
typedef struct { int32_t d, magic, shift, add_sign; } pjo_fastdiv;
void    pjo_fastdiv_init(pjo_fastdiv* f, int32_t d);
int32_t pjo_fastdiv_div(int32_t n, const pjo_fastdiv* f);
int32_t pjo_fastdiv_mod(int32_t n, const pjo_fastdiv* f);
/* Slot index the join kernel uses: |(int)h % size| (src/hash_join_kernels.cuh:100-101). */
int32_t pjo_hash_slot(uint32_t h, int32_t table_size);

/* ---- Seeded generators (own definitions; the reference's cuRAND stream is
 *      not reproducible).  Semantics follow src/generate_input_tables.cuh:43-114
 *      (unique lottery, selectivity) and src/generate_zipf_input.cuh:50-127
 *      (Zipf CDF table + binary search). ---------------------------------- */
uint64_t pjo_mix64(uint64_t z);                        /* SplitMix64 finaliser */
uint64_t pjo_rand64(uint64_t seed, uint64_t i);        /* i-th SplitMix64 output */
uint32_t pjo_bounded(uint64_t r, uint64_t n);          /* floor(hi32(r) * n / 2^32), n <= 2^32 */
uint64_t pjo_perm(uint64_t i, uint64_t n, uint64_t seed); /* keyed bijection on [0,n) */
/* keys[i] = perm_n(i) for i < count (unique build keys drawn from [0,domain)). */
void pjo_gen_build_unique_i32(int32_t* keys, int64_t count, int64_t domain, uint64_t perm_seed);
/* keys[i] = perm_domain(U(0,m)) : build keys with repeats. */
void pjo_gen_build_dup_i32(int32_t* keys, int64_t count, int64_t m, int64_t domain,
                           uint64_t perm_seed, uint64_t seed);
/* Probe keys: with probability sel a key of the build set perm_domain(U(0,n_build)),
 * otherwise, lottery mode (nonmatch_range == 0): perm_domain(n_build + U(0, domain-n_build)),
 * else nonmatch_base + U(0, nonmatch_range). */
void pjo_gen_probe_i32(int32_t* keys, int64_t count, int64_t domain, int64_t n_build,
                       uint64_t perm_seed, uint64_t seed, double sel,
                       int64_t nonmatch_base, int64_t nonmatch_range);
/* Zipf CDF table: lut[i-1] = sum_{j<=i} j^-theta / sum_{j<=n} j^-theta. */
void pjo_zipf_lut(double* lut, int64_t n, double theta);
/* Smallest index with lut[idx] >= u, clamped to n-1 (src/generate_zipf_input.cuh:104-124). */
int64_t pjo_zipf_rank(const double* lut, int64_t n, double u);
/* Probe keys = perm_{n}(rank) with rank ~ Zipf(theta) over [0,n) and the
 * uniform u = (rand64(seed,j) >> 11) * 2^-53. */
void pjo_gen_probe_zipf_i32(int32_t* keys, int64_t count, const double* lut, int64_t n,
                            uint64_t rank_perm_seed, uint64_t seed);
/* 64-bit keys: build key i = rand64(key_seed, i) (unique: mix64 is a bijection). */
void pjo_gen_build_i64(int64_t* keys, int64_t count, uint64_t key_seed);
void pjo_gen_probe_i64(int64_t* keys, int64_t count, int64_t n_build, uint64_t key_seed,
                       uint64_t seed, double sel);
void pjo_gen_probe_zipf_i64(int64_t* keys, int64_t count, const double* lut, int64_t n_build,
                            uint64_t key_seed, uint64_t rank_perm_seed, uint64_t seed);

/* ---- Joins (src/inner_join_cpu.hpp:23-107) ------------------------------- */
/* Restatement of inner_join_cpu: std::unordered_map (unique keys) or
 * std::unordered_multimap reserved to 2|R|, serial build, OpenMP-parallel probe
 * with an atomic output cursor for unique keys, serial probe otherwise.
 * Emits (build payload, probe payload); payload NULL means row id.
 * Returns the number of pairs found; writes at most cap pairs.
 * *seconds receives the build+probe wall time (omp_get_wtime window,
 * src/inner_join_cpu.hpp:48,77). key_bytes is 4 or 8. */
int64_t pjo_inner_join_cpu(const void* build_keys, const int32_t* build_pay, int64_t nb,
                           const void* probe_keys, const int32_t* probe_pay, int64_t np,
                           int key_bytes, int uniq_keys, int num_threads,
                           int32_t* out_pairs, int64_t cap, double* seconds);
/* Independent sort-merge join (std::sort both sides, merge); same output contract. */
int64_t pjo_sort_merge_join(const void* build_keys, const int32_t* build_pay, int64_t nb,
                            const void* probe_keys, const int32_t* probe_pay, int64_t np,
                            int key_bytes, int32_t* out_pairs, int64_t cap);
/* |R join S| = sum_k cnt_R(k) * cnt_S(k), by sorting (64-bit count). */
int64_t pjo_join_count(const void* build_keys, int64_t nb, const void* probe_keys, int64_t np,
                       int key_bytes);
/* Sort pairs lexicographically (first, second) in place (src/inner_join_cpu.hpp:120-134
 * sorts by .x then stably by .y; comparing sorted multisets is what matters). */
void pjo_sort_pairs(int32_t* pairs, int64_t n);

/* ---- Partition phases ---------------------------------------------------- */
/* digit = (key >> shift) & (2^bits - 1) over the low 32 bits of the key
 * (src/partition_kernels.cuh:109,168,177,380,389). hist has 2^bits entries. */
void pjo_histogram(const void* keys, int64_t n, int key_bytes, int shift, int bits, int64_t* hist);
/* One scatter pass restated as a stable counting sort by digit within each
 * input segment; the reference's intra-partition order is unspecified, so
 * tests compare it as a per-partition multiset. */
void pjo_partition_pass(const void* keys_in, const int32_t* pay_in, int64_t n, int key_bytes,
                        int shift, int bits, void* keys_out, int32_t* pay_out, int64_t* offsets);

#ifdef __cplusplus
}
#endif
#endif
