/*
 * ref_shim.cpp -- extern "C" entry points over the reference's own CPU code.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) only where
 * the reference sources are present, with -I pointing at the reference's src/
 * directory; the library lands in oracle/_ref/ (git-ignored).  It compiles the
 * reference's own headers unchanged:
 *   src/hash_functions.cuh  MurmurHash3_32<Key>          (lines 33-125)
 *   src/int_fastdiv.h       int_fastdiv, operator/ and % (lines 21-133)
 *   src/inner_join_cpu.hpp  inner_join_cpu                (lines 23-107)
 * and exposes them so tests can pin the oracle restatement against them and
 * bench.py can time the reference CPU join (cpu_baseline kind "reference").
 */
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <limits>
#include <vector>

#include "hash_functions.cuh"
#include "common.cuh"
#include "int_fastdiv.h"
#include "inner_join_cpu.hpp"

extern "C" uint32_t pjref_murmur3_i32(int32_t key) {
  MurmurHash3_32<int> h;
  return h(int(key));
}

extern "C" uint32_t pjref_murmur3_i64(int64_t key) {
  MurmurHash3_32<long long> h;
  return h((long long)key);
}

extern "C" int32_t pjref_fastdiv_div(int32_t n, int32_t d) {
  int_fastdiv f(d);
  return n / f;
}

extern "C" int32_t pjref_fastdiv_mod(int32_t n, int32_t d) {
  int_fastdiv f(d);
  return n % f;
}

/* Runs the reference inner_join_cpu on copies of the keys.  Writes (x, y) of
 * each joined_type element (x = key, y = 0 in the reference) into out_xy, up to
 * cap pairs.  Returns the pair count; *seconds = the reference's own timing. */
template <typename K, typename J>
static int64_t run_ref_join(const K* build, int64_t nb, const K* probe, int64_t np, int uniq,
                            K* out_xy, int64_t cap, int nthreads, double* seconds) {
  std::vector<K> b(build, build + nb), p(probe, probe + np);
  std::vector<J> joined;
  joined.reserve(size_t(cap > 0 ? cap : 1));
  double t = inner_join_cpu<K, J>(b, uniq != 0, p, joined, nthreads);
  if (seconds) *seconds = t;
  const int64_t n = int64_t(joined.size());
  for (int64_t i = 0; i < n && i < cap; ++i) {
    out_xy[2 * i] = joined[size_t(i)].x;
    out_xy[2 * i + 1] = joined[size_t(i)].y;
  }
  return n;
}

extern "C" int64_t pjref_inner_join_cpu_i32(const int32_t* build, int64_t nb, const int32_t* probe, int64_t np,
                                            int uniq, int32_t* out_xy, int64_t cap, int nthreads,
                                            double* seconds) {
  return run_ref_join<int, int2>(build, nb, probe, np, uniq, out_xy, cap, nthreads, seconds);
}

extern "C" int64_t pjref_inner_join_cpu_i64(const int64_t* build, int64_t nb, const int64_t* probe, int64_t np,
                                            int uniq, int64_t* out_xy, int64_t cap, int nthreads,
                                            double* seconds) {
  return run_ref_join<long long, longlong2>(reinterpret_cast<const long long*>(build), nb,
                                            reinterpret_cast<const long long*>(probe), np, uniq,
                                            reinterpret_cast<long long*>(out_xy), cap, nthreads, seconds);
}
