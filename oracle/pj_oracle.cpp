/*
 * pj_oracle.cpp -- CPU restatement of the reference join path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  The product library never links or
 * loads this file's object code; see pj_oracle.h for the list of functions and
 * the reference lines each one restates.
 */
#include "pj_oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <utility>
#include <vector>

/* ------------------------------------------------------------------------ */
/* MurmurHash3_x86_32.  Restates src/hash_functions.cuh:80-119 (body, tail,
 * finaliser) for arbitrary byte strings so that the SMHasher verification
 * procedure can pin it.                                                     */
/* ------------------------------------------------------------------------ */
static inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

static inline uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

extern "C" uint32_t pjo_murmur3_32(const void* data, int len, uint32_t seed) {
  const uint8_t* bytes = static_cast<const uint8_t*>(data);
  const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
  uint32_t h = seed;
  const int nblocks = len / 4;
  for (int b = 0; b < nblocks; ++b) {
    uint32_t k;
    std::memcpy(&k, bytes + 4 * b, 4);
    k *= c1;
    k = rotl32(k, 15);
    k *= c2;
    h ^= k;
    h = rotl32(h, 13);
    h = h * 5 + 0xe6546b64u;
  }
  const uint8_t* tail = bytes + 4 * nblocks;
  uint32_t k = 0;
  switch (len & 3) {
    case 3: k ^= uint32_t(tail[2]) << 16; /* fall through */
    case 2: k ^= uint32_t(tail[1]) << 8;  /* fall through */
    case 1:
      k ^= tail[0];
      k *= c1;
      k = rotl32(k, 15);
      k *= c2;
      h ^= k;
  }
  h ^= uint32_t(len);
  return fmix32(h);
}

extern "C" uint32_t pjo_murmur3_key4(uint32_t key) { return pjo_murmur3_32(&key, 4, 0); }
extern "C" uint32_t pjo_murmur3_key8(uint64_t key) { return pjo_murmur3_32(&key, 8, 0); }

/* ------------------------------------------------------------------------ */
/* int_fastdiv: Hacker's Delight 2nd ed. section 10-4 "magic" for signed
 * divisors (the algorithm src/int_fastdiv.h:44-96 implements), plus the
 * d = +-1 special case (src/int_fastdiv.h:47-57).                          */
/* ------------------------------------------------------------------------ */
extern "C" void pjo_fastdiv_init(pjo_fastdiv* f, int32_t d) {
  f->d = d;
  if (d == 1 || d == -1) {
    f->magic = 0;
    f->shift = -1;
    f->add_sign = d;
    return;
  }
  if (d == 0) { /* undefined; keep the struct inert */
    f->magic = 0;
    f->shift = -1;
    f->add_sign = 0;
    return;
  }
  const uint32_t two31 = 0x80000000u;
  const uint32_t ad = d < 0 ? uint32_t(0) - uint32_t(d) : uint32_t(d);
  const uint32_t t = two31 + (uint32_t(d) >> 31);
  const uint32_t anc = t - 1 - t % ad;
  int p = 31;
  uint32_t q1 = two31 / anc, r1 = two31 - q1 * anc;
  uint32_t q2 = two31 / ad, r2 = two31 - q2 * ad;
  uint32_t delta;
  do {
    ++p;
    q1 <<= 1; r1 <<= 1;
    if (r1 >= anc) { ++q1; r1 -= anc; }
    q2 <<= 1; r2 <<= 1;
    if (r2 >= ad) { ++q2; r2 -= ad; }
    delta = ad - r2;
  } while (q1 < delta || (q1 == delta && r1 == 0));
  int32_t m = int32_t(q2 + 1);
  if (d < 0) m = -m;
  f->magic = m;
  f->shift = p - 32;
  f->add_sign = (d > 0 && m < 0) ? 1 : ((d < 0 && m > 0) ? -1 : 0);
}

extern "C" int32_t pjo_fastdiv_div(int32_t n, const pjo_fastdiv* f) {
  int32_t q = int32_t((int64_t(f->magic) * int64_t(n)) >> 32);
  q += n * f->add_sign;
  if (f->shift >= 0) {
    q >>= f->shift;
    q += int32_t(uint32_t(q) >> 31);
  }
  return q;
}

extern "C" int32_t pjo_fastdiv_mod(int32_t n, const pjo_fastdiv* f) {
  return n - pjo_fastdiv_div(n, f) * f->d;
}

extern "C" int32_t pjo_hash_slot(uint32_t h, int32_t table_size) {
  pjo_fastdiv f;
  pjo_fastdiv_init(&f, table_size);
  int32_t r = pjo_fastdiv_mod(int32_t(h), &f);
  return r < 0 ? -r : r;
}

/* ------------------------------------------------------------------------ */
/* Generators.  SplitMix64 (Steele, Lea, Flood, OOPSLA 2014) as a counter-
 * based stream; a keyed bijection on [0, 2^k) with cycle walking for
 * permutations.                                                             */
/* ------------------------------------------------------------------------ */
static const uint64_t kGolden = 0x9e3779b97f4a7c15ull;

extern "C" uint64_t pjo_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

extern "C" uint64_t pjo_rand64(uint64_t seed, uint64_t i) { return pjo_mix64(seed + (i + 1) * kGolden); }

extern "C" uint32_t pjo_bounded(uint64_t r, uint64_t n) { return uint32_t(((r >> 32) * n) >> 32); }

static inline int ceil_log2(uint64_t n) {
  int k = 0;
  while ((uint64_t(1) << k) < n) ++k;
  return k < 1 ? 1 : k;
}

/* Four rounds of (x*a + b mod 2^k; x ^= x >> s) with odd a: each step is a
 * bijection of [0, 2^k), so the composition is too. */
static inline uint64_t perm_round_fn(uint64_t x, int k, const uint64_t* a, const uint64_t* b) {
  const uint64_t mask = (k >= 64) ? ~0ull : ((uint64_t(1) << k) - 1);
  const int s = (k + 1) / 2;
  for (int r = 0; r < 4; ++r) {
    x = (x * a[r] + b[r]) & mask;
    x ^= x >> s;
  }
  return x;
}

static inline void perm_keys(uint64_t seed, uint64_t* a, uint64_t* b) {
  for (int r = 0; r < 4; ++r) {
    a[r] = pjo_rand64(seed, 2 * r) | 1ull;
    b[r] = pjo_rand64(seed, 2 * r + 1);
  }
}

extern "C" uint64_t pjo_perm(uint64_t i, uint64_t n, uint64_t seed) {
  uint64_t a[4], b[4];
  perm_keys(seed, a, b);
  const int k = ceil_log2(n);
  uint64_t x = perm_round_fn(i, k, a, b);
  while (x >= n) x = perm_round_fn(x, k, a, b);
  return x;
}

extern "C" void pjo_gen_build_unique_i32(int32_t* keys, int64_t count, int64_t domain, uint64_t perm_seed) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; ++i) keys[i] = int32_t(pjo_perm(uint64_t(i), uint64_t(domain), perm_seed));
}

extern "C" void pjo_gen_build_dup_i32(int32_t* keys, int64_t count, int64_t m, int64_t domain,
                                      uint64_t perm_seed, uint64_t seed) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; ++i) {
    uint64_t u = pjo_bounded(pjo_rand64(seed, uint64_t(i)), uint64_t(m));
    keys[i] = int32_t(pjo_perm(u, uint64_t(domain), perm_seed));
  }
}

static inline uint64_t sel_threshold(double sel) {
  if (sel >= 1.0) return uint64_t(1) << 32;
  if (sel <= 0.0) return 0;
  return uint64_t(sel * 4294967296.0);
}

extern "C" void pjo_gen_probe_i32(int32_t* keys, int64_t count, int64_t domain, int64_t n_build,
                                  uint64_t perm_seed, uint64_t seed, double sel,
                                  int64_t nonmatch_base, int64_t nonmatch_range) {
  const uint64_t thr = sel_threshold(sel);
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < count; ++j) {
    const uint64_t r1 = pjo_rand64(seed, 2 * uint64_t(j));
    const uint64_t r2 = pjo_rand64(seed, 2 * uint64_t(j) + 1);
    int64_t key;
    if ((r1 >> 32) < thr) {
      key = int64_t(pjo_perm(pjo_bounded(r2, uint64_t(n_build)), uint64_t(domain), perm_seed));
    } else if (nonmatch_range == 0) {
      key = int64_t(pjo_perm(uint64_t(n_build) + pjo_bounded(r2, uint64_t(domain - n_build)),
                             uint64_t(domain), perm_seed));
    } else {
      key = nonmatch_base + int64_t(pjo_bounded(r2, uint64_t(nonmatch_range)));
    }
    keys[j] = int32_t(key);
  }
}

extern "C" void pjo_zipf_lut(double* lut, int64_t n, double theta) {
  double scale = 0.0;
  for (int64_t i = 1; i <= n; ++i) scale += 1.0 / std::pow(double(i), theta);
  double sum = 0.0;
  for (int64_t i = 1; i <= n; ++i) {
    sum += 1.0 / std::pow(double(i), theta);
    lut[i - 1] = sum / scale;
  }
}

extern "C" int64_t pjo_zipf_rank(const double* lut, int64_t n, double u) {
  if (lut[0] >= u) return 0;
  int64_t lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (lut[mid] < u) lo = mid; else hi = mid;
  }
  return hi;
}

static inline double unit_double(uint64_t r) { return double(r >> 11) * (1.0 / 9007199254740992.0); }

extern "C" void pjo_gen_probe_zipf_i32(int32_t* keys, int64_t count, const double* lut, int64_t n,
                                       uint64_t rank_perm_seed, uint64_t seed) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < count; ++j) {
    const int64_t rank = pjo_zipf_rank(lut, n, unit_double(pjo_rand64(seed, uint64_t(j))));
    keys[j] = int32_t(pjo_perm(uint64_t(rank), uint64_t(n), rank_perm_seed));
  }
}

extern "C" void pjo_gen_build_i64(int64_t* keys, int64_t count, uint64_t key_seed) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; ++i) keys[i] = int64_t(pjo_rand64(key_seed, uint64_t(i)));
}

extern "C" void pjo_gen_probe_i64(int64_t* keys, int64_t count, int64_t n_build, uint64_t key_seed,
                                  uint64_t seed, double sel) {
  const uint64_t thr = sel_threshold(sel);
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < count; ++j) {
    const uint64_t r1 = pjo_rand64(seed, 2 * uint64_t(j));
    const uint64_t r2 = pjo_rand64(seed, 2 * uint64_t(j) + 1);
    uint64_t idx;
    if ((r1 >> 32) < thr) idx = pjo_bounded(r2, uint64_t(n_build));
    else idx = uint64_t(n_build) + (r2 >> 2);  /* counters past the build set: never a build key */
    keys[j] = int64_t(pjo_rand64(key_seed, idx));
  }
}

extern "C" void pjo_gen_probe_zipf_i64(int64_t* keys, int64_t count, const double* lut, int64_t n_build,
                                       uint64_t key_seed, uint64_t rank_perm_seed, uint64_t seed) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < count; ++j) {
    const int64_t rank = pjo_zipf_rank(lut, n_build, unit_double(pjo_rand64(seed, uint64_t(j))));
    const uint64_t idx = pjo_perm(uint64_t(rank), uint64_t(n_build), rank_perm_seed);
    keys[j] = int64_t(pjo_rand64(key_seed, idx));
  }
}

/* ------------------------------------------------------------------------ */
/* Joins                                                                     */
/* ------------------------------------------------------------------------ */
template <typename K>
static inline K key_at(const void* keys, int64_t i) { return static_cast<const K*>(keys)[i]; }

static inline int32_t pay_at(const int32_t* pay, int64_t i) { return pay ? pay[i] : int32_t(i); }

/* Restates src/inner_join_cpu.hpp:39-104. */
template <typename K>
static int64_t inner_join_cpu_t(const void* bk, const int32_t* bp, int64_t nb, const void* pk,
                                const int32_t* pp, int64_t np, int uniq, int32_t* out, int64_t cap,
                                double* seconds) {
  double t0 = 0, t1 = 0;
  int64_t found = 0;
  if (uniq) {
    std::unordered_map<K, int64_t> ht;
    ht.reserve(size_t(2 * nb));
    const auto end = ht.end();
    t0 = omp_get_wtime();
    for (int64_t i = 0; i < nb; ++i) ht.insert(std::make_pair(key_at<K>(bk, i), i));
    std::atomic<int64_t> cursor{0};
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < np; ++j) {
      auto it = ht.find(key_at<K>(pk, j));
      if (it != end) {
        const int64_t slot = cursor.fetch_add(1, std::memory_order_relaxed);
        if (slot < cap) {
          out[2 * slot] = pay_at(bp, it->second);
          out[2 * slot + 1] = pay_at(pp, j);
        }
      }
    }
    t1 = omp_get_wtime();
    found = cursor.load();
  } else {
    std::unordered_multimap<K, int64_t> ht;
    ht.reserve(size_t(2 * nb));
    t0 = omp_get_wtime();
    for (int64_t i = 0; i < nb; ++i) ht.insert(std::make_pair(key_at<K>(bk, i), i));
    for (int64_t j = 0; j < np; ++j) {
      auto range = ht.equal_range(key_at<K>(pk, j));
      for (auto it = range.first; it != range.second; ++it) {
        if (found < cap) {
          out[2 * found] = pay_at(bp, it->second);
          out[2 * found + 1] = pay_at(pp, j);
        }
        ++found;
      }
    }
    t1 = omp_get_wtime();
  }
  if (seconds) *seconds = t1 - t0;
  return found;
}

extern "C" int64_t pjo_inner_join_cpu(const void* build_keys, const int32_t* build_pay, int64_t nb,
                                      const void* probe_keys, const int32_t* probe_pay, int64_t np,
                                      int key_bytes, int uniq_keys, int num_threads, int32_t* out_pairs,
                                      int64_t cap, double* seconds) {
  if (num_threads > 0) omp_set_num_threads(num_threads);
  if (key_bytes == 8)
    return inner_join_cpu_t<int64_t>(build_keys, build_pay, nb, probe_keys, probe_pay, np, uniq_keys,
                                     out_pairs, cap, seconds);
  return inner_join_cpu_t<int32_t>(build_keys, build_pay, nb, probe_keys, probe_pay, np, uniq_keys,
                                   out_pairs, cap, seconds);
}

template <typename K>
static int64_t sort_merge_t(const void* bk, const int32_t* bp, int64_t nb, const void* pk,
                            const int32_t* pp, int64_t np, int32_t* out, int64_t cap) {
  std::vector<std::pair<K, int64_t>> r(static_cast<size_t>(nb));
  std::vector<std::pair<K, int64_t>> s(static_cast<size_t>(np));
  for (int64_t i = 0; i < nb; ++i) r[size_t(i)] = {key_at<K>(bk, i), i};
  for (int64_t j = 0; j < np; ++j) s[size_t(j)] = {key_at<K>(pk, j), j};
  std::sort(r.begin(), r.end());
  std::sort(s.begin(), s.end());
  int64_t found = 0;
  size_t i = 0, j = 0;
  while (i < r.size() && j < s.size()) {
    if (r[i].first < s[j].first) { ++i; continue; }
    if (s[j].first < r[i].first) { ++j; continue; }
    size_t i2 = i, j2 = j;
    while (i2 < r.size() && r[i2].first == r[i].first) ++i2;
    while (j2 < s.size() && s[j2].first == s[j].first) ++j2;
    for (size_t a = i; a < i2; ++a)
      for (size_t b = j; b < j2; ++b) {
        if (found < cap) {
          out[2 * found] = pay_at(bp, r[a].second);
          out[2 * found + 1] = pay_at(pp, s[b].second);
        }
        ++found;
      }
    i = i2;
    j = j2;
  }
  return found;
}

extern "C" int64_t pjo_sort_merge_join(const void* build_keys, const int32_t* build_pay, int64_t nb,
                                       const void* probe_keys, const int32_t* probe_pay, int64_t np,
                                       int key_bytes, int32_t* out_pairs, int64_t cap) {
  if (key_bytes == 8)
    return sort_merge_t<int64_t>(build_keys, build_pay, nb, probe_keys, probe_pay, np, out_pairs, cap);
  return sort_merge_t<int32_t>(build_keys, build_pay, nb, probe_keys, probe_pay, np, out_pairs, cap);
}

template <typename K>
static int64_t join_count_t(const void* bk, int64_t nb, const void* pk, int64_t np) {
  std::vector<K> r(static_cast<const K*>(bk), static_cast<const K*>(bk) + nb);
  std::vector<K> s(static_cast<const K*>(pk), static_cast<const K*>(pk) + np);
  std::sort(r.begin(), r.end());
  std::sort(s.begin(), s.end());
  int64_t total = 0;
  size_t i = 0, j = 0;
  while (i < r.size() && j < s.size()) {
    if (r[i] < s[j]) { ++i; continue; }
    if (s[j] < r[i]) { ++j; continue; }
    size_t i2 = i, j2 = j;
    while (i2 < r.size() && r[i2] == r[i]) ++i2;
    while (j2 < s.size() && s[j2] == s[j]) ++j2;
    total += int64_t(i2 - i) * int64_t(j2 - j);
    i = i2;
    j = j2;
  }
  return total;
}

extern "C" int64_t pjo_join_count(const void* build_keys, int64_t nb, const void* probe_keys, int64_t np,
                                  int key_bytes) {
  if (key_bytes == 8) return join_count_t<int64_t>(build_keys, nb, probe_keys, np);
  return join_count_t<int32_t>(build_keys, nb, probe_keys, np);
}

extern "C" void pjo_sort_pairs(int32_t* pairs, int64_t n) {
  uint64_t* p = reinterpret_cast<uint64_t*>(pairs);
  /* order by (first, second): pack as (uint32(first) ^ sign) << 32 | (uint32(second) ^ sign) */
  std::vector<uint64_t> tmp(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t a = uint32_t(pairs[2 * i]) ^ 0x80000000u, b = uint32_t(pairs[2 * i + 1]) ^ 0x80000000u;
    tmp[size_t(i)] = (uint64_t(a) << 32) | b;
  }
  std::sort(tmp.begin(), tmp.end());
  for (int64_t i = 0; i < n; ++i) {
    pairs[2 * i] = int32_t(uint32_t(tmp[size_t(i)] >> 32) ^ 0x80000000u);
    pairs[2 * i + 1] = int32_t(uint32_t(tmp[size_t(i)]) ^ 0x80000000u);
  }
  (void)p;
}

/* ------------------------------------------------------------------------ */
/* Partition phases                                                          */
/* ------------------------------------------------------------------------ */
static inline uint32_t digit_of(const void* keys, int key_bytes, int64_t i, int shift, int bits) {
  const uint64_t k = key_bytes == 8 ? uint64_t(static_cast<const int64_t*>(keys)[i])
                                    : uint64_t(uint32_t(static_cast<const int32_t*>(keys)[i]));
  return uint32_t((k >> shift) & ((uint64_t(1) << bits) - 1));
}

extern "C" void pjo_histogram(const void* keys, int64_t n, int key_bytes, int shift, int bits, int64_t* hist) {
  const int64_t nb = int64_t(1) << bits;
  std::fill(hist, hist + nb, 0);
  for (int64_t i = 0; i < n; ++i) ++hist[digit_of(keys, key_bytes, i, shift, bits)];
}

extern "C" void pjo_partition_pass(const void* keys_in, const int32_t* pay_in, int64_t n, int key_bytes,
                                   int shift, int bits, void* keys_out, int32_t* pay_out, int64_t* offsets) {
  const int64_t nb = int64_t(1) << bits;
  std::vector<int64_t> hist(static_cast<size_t>(nb), 0);
  pjo_histogram(keys_in, n, key_bytes, shift, bits, hist.data());
  offsets[0] = 0;
  for (int64_t b = 0; b < nb; ++b) offsets[b + 1] = offsets[b] + hist[size_t(b)];
  std::vector<int64_t> cur(offsets, offsets + nb);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t pos = cur[digit_of(keys_in, key_bytes, i, shift, bits)]++;
    if (key_bytes == 8)
      static_cast<int64_t*>(keys_out)[pos] = static_cast<const int64_t*>(keys_in)[i];
    else
      static_cast<int32_t*>(keys_out)[pos] = static_cast<const int32_t*>(keys_in)[i];
    pay_out[pos] = pay_at(pay_in, i);
  }
}
