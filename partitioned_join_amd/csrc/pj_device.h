// pj_device.h -- device helpers for the gfx950 radix-partitioned join kernels.
//
// Tuple layouts, MurmurHash3_x86_32, int_fastdiv, wave64 ballot helpers, block
// scans and the SplitMix64 / keyed-bijection generator arithmetic.  Integer-only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pj {

constexpr int kNumXcd = 8;     // MI355X: 8 XCDs; workgroup b runs on XCD b % 8
constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Tuple layouts.  4-byte keys: one 8-byte word {key, payload}.  8-byte keys:
// 16 bytes {key lo, key hi, payload, 0}.
// ---------------------------------------------------------------------------
template <int KB> struct KeyT;

template <> struct KeyT<4> {
  using key_t = uint32_t;
  using tup_t = uint2;
  __device__ __forceinline__ static key_t key(const tup_t& t) { return t.x; }
  __device__ __forceinline__ static int32_t pay(const tup_t& t) { return int32_t(t.y); }
  __device__ __forceinline__ static tup_t make(key_t k, int32_t p) { return make_uint2(k, uint32_t(p)); }
  __device__ __forceinline__ static uint32_t low32(key_t k) { return k; }
};

// A plain 16-byte aggregate rather than HIP's uint4 (a union-based vector
// type): arrays of it held in registers are split into scalars by the
// compiler, where arrays of uint4 can end up in scratch memory.
struct alignas(16) Tup8 {
  uint32_t x, y, z, w;
};

template <> struct KeyT<8> {
  using key_t = uint64_t;
  using tup_t = Tup8;
  __device__ __forceinline__ static key_t key(const tup_t& t) { return uint64_t(t.x) | (uint64_t(t.y) << 32); }
  __device__ __forceinline__ static int32_t pay(const tup_t& t) { return int32_t(t.z); }
  __device__ __forceinline__ static tup_t make(key_t k, int32_t p) {
    return Tup8{uint32_t(k), uint32_t(k >> 32), uint32_t(p), 0u};
  }
  __device__ __forceinline__ static uint32_t low32(key_t k) { return uint32_t(k); }
};

// Radix digit over the key's low 32 bits (reference: src/partition_kernels.cuh:109,168,380).
__device__ __forceinline__ uint32_t digit_of(uint32_t low32, int shift, uint32_t mask) {
  return (low32 >> shift) & mask;
}

// ---------------------------------------------------------------------------
// MurmurHash3_x86_32 with seed 0 over the key's little-endian bytes
// (reference: src/hash_functions.cuh:80-119; Appleby, public domain).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

__device__ __forceinline__ uint32_t murmur_block(uint32_t h, uint32_t k) {
  k *= 0xcc9e2d51u;
  k = rotl32(k, 15);
  k *= 0x1b873593u;
  h ^= k;
  h = rotl32(h, 13);
  return h * 5u + 0xe6546b64u;
}

__device__ __forceinline__ uint32_t murmur_fmix(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

__device__ __forceinline__ uint32_t murmur3(uint32_t key) { return murmur_fmix(murmur_block(0u, key) ^ 4u); }

__device__ __forceinline__ uint32_t murmur3(uint64_t key) {
  uint32_t h = murmur_block(0u, uint32_t(key));
  h = murmur_block(h, uint32_t(key >> 32));
  return murmur_fmix(h ^ 8u);
}

// ---------------------------------------------------------------------------
// int_fastdiv: signed division by a runtime-invariant divisor using the magic
// number and shift of Hacker's Delight (2nd ed.) section 10-4 (reference:
// src/int_fastdiv.h:21-133).  Host and device share the set-up routine.
// ---------------------------------------------------------------------------
struct FastDiv {
  int32_t d, magic, shift, add_sign;
};

__host__ __device__ inline FastDiv fastdiv_make(int32_t d) {
  FastDiv f;
  f.d = d;
  if (d == 1 || d == -1 || d == 0) {
    f.magic = 0;
    f.shift = -1;
    f.add_sign = (d == 0) ? 0 : d;
    return f;
  }
  const uint32_t two31 = 0x80000000u;
  const uint32_t ad = d < 0 ? 0u - uint32_t(d) : uint32_t(d);
  const uint32_t t = two31 + (uint32_t(d) >> 31);
  const uint32_t anc = t - 1u - t % ad;
  int p = 31;
  uint32_t q1 = two31 / anc, r1 = two31 - q1 * anc;
  uint32_t q2 = two31 / ad, r2 = two31 - q2 * ad;
  uint32_t delta;
  do {
    ++p;
    q1 <<= 1;
    r1 <<= 1;
    if (r1 >= anc) { ++q1; r1 -= anc; }
    q2 <<= 1;
    r2 <<= 1;
    if (r2 >= ad) { ++q2; r2 -= ad; }
    delta = ad - r2;
  } while (q1 < delta || (q1 == delta && r1 == 0));
  int32_t m = int32_t(q2 + 1u);
  if (d < 0) m = -m;
  f.magic = m;
  f.shift = p - 32;
  f.add_sign = (d > 0 && m < 0) ? 1 : ((d < 0 && m > 0) ? -1 : 0);
  return f;
}

__device__ __forceinline__ int32_t fastdiv_div(int32_t n, const FastDiv& f) {
  int32_t q = __mulhi(f.magic, n);
  q += n * f.add_sign;
  if (f.shift >= 0) {
    q >>= f.shift;
    q += int32_t(uint32_t(q) >> 31);
  }
  return q;
}

__device__ __forceinline__ int32_t fastdiv_mod(int32_t n, const FastDiv& f) { return n - fastdiv_div(n, f) * f.d; }

// Slot of a hash in a table of f.d slots: |(int)h % size| (src/hash_join_kernels.cuh:100-101).
__device__ __forceinline__ uint32_t hash_slot(uint32_t h, const FastDiv& f) {
  const int32_t r = fastdiv_mod(int32_t(h), f);
  return uint32_t(r < 0 ? -r : r);
}

// ---------------------------------------------------------------------------
// Barriers.  __syncthreads() carries a workgroup-scope fence, which waits for
// every outstanding global load and store of the wave (vmcnt(0)) before the
// barrier.  Where a barrier only orders LDS data, lds_barrier() waits for this
// wave's LDS operations only, so prefetched global loads and write-backs stay
// in flight across it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void lds_barrier() { __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Block-wide OR of `pred` with one LDS barrier and no global fence.  `slots`
// is 3 LDS words zeroed before first use; `round` counts calls (same on every
// thread).
__device__ __forceinline__ bool block_any(bool pred, uint32_t* slots, uint32_t round) {
  const uint32_t cur = round % 3, nxt = (round + 1) % 3;
  if (threadIdx.x == 0) slots[nxt] = 0;  // last read before the previous barrier
  if (__ballot(pred) && (threadIdx.x & (kWave - 1)) == 0) slots[cur] = 1;
  lds_barrier();
  return slots[cur] != 0;
}

// ---------------------------------------------------------------------------
// wave64 helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Number of set bits of `mask` below this lane (v_mbcnt).
__device__ __forceinline__ uint32_t mbcnt(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// Block-wide exclusive scan of one uint32 per thread.  s_tmp needs BLOCK/64 + 1
// entries.  Returns the exclusive prefix; *total receives the block sum.
// Contains two LDS barriers.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
  constexpr int NW = BLOCK / kWave;
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, kWave);
    if (lane >= uint32_t(d)) x += y;
  }
  if (lane == kWave - 1) s_tmp[wid] = x;
  lds_barrier();
  if (wid == 0) {
    const uint32_t w = lane < uint32_t(NW) ? s_tmp[lane] : 0u;
    uint32_t wx = w;
#pragma unroll
    for (int d = 1; d < NW; d <<= 1) {
      const uint32_t y = __shfl_up(wx, d, kWave);
      if (lane >= uint32_t(d)) wx += y;
    }
    if (lane < uint32_t(NW)) s_tmp[lane] = wx - w;
    if (lane == uint32_t(NW - 1)) s_tmp[NW] = wx;
  }
  lds_barrier();
  const uint32_t res = x - v + s_tmp[wid];
  *total = s_tmp[NW];
  return res;
}

// 64-bit variant.
template <int BLOCK>
__device__ __forceinline__ uint64_t block_exclusive_scan64(uint64_t v, uint64_t* s_tmp, uint64_t* total) {
  constexpr int NW = BLOCK / kWave;
  const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  uint64_t x = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint64_t y = __shfl_up(x, d, kWave);
    if (lane >= uint32_t(d)) x += y;
  }
  if (lane == kWave - 1) s_tmp[wid] = x;
  lds_barrier();
  if (wid == 0) {
    const uint64_t w = lane < uint32_t(NW) ? s_tmp[lane] : 0ull;
    uint64_t wx = w;
#pragma unroll
    for (int d = 1; d < NW; d <<= 1) {
      const uint64_t y = __shfl_up(wx, d, kWave);
      if (lane >= uint32_t(d)) wx += y;
    }
    if (lane < uint32_t(NW)) s_tmp[lane] = wx - w;
    if (lane == uint32_t(NW - 1)) s_tmp[NW] = wx;
  }
  lds_barrier();
  const uint64_t res = x - v + s_tmp[wid];
  *total = s_tmp[NW];
  return res;
}

// ---------------------------------------------------------------------------
// Generators: SplitMix64 (Steele, Lea, Flood 2014) counters and a keyed
// bijection of [0, 2^k) (four rounds of odd multiply-add and xorshift) with
// cycle walking onto [0, n).
// ---------------------------------------------------------------------------
constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ull;

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t rand64(uint64_t seed, uint64_t i) {
  return mix64(seed + (i + 1) * kGolden);
}

__host__ __device__ __forceinline__ uint32_t bounded(uint64_t r, uint64_t n) {
  return uint32_t(((r >> 32) * n) >> 32);
}

struct PermKey {
  uint64_t a[4], b[4];
  uint64_t n, mask;
  int k, s;
};

__host__ __device__ inline PermKey perm_key(uint64_t n, uint64_t seed) {
  PermKey p;
  for (int r = 0; r < 4; ++r) {
    p.a[r] = rand64(seed, 2 * r) | 1ull;
    p.b[r] = rand64(seed, 2 * r + 1);
  }
  int k = 0;
  while ((uint64_t(1) << k) < n) ++k;
  if (k < 1) k = 1;
  p.k = k;
  p.s = (k + 1) / 2;
  p.mask = (k >= 64) ? ~0ull : ((uint64_t(1) << k) - 1);
  p.n = n;
  return p;
}

__host__ __device__ __forceinline__ uint64_t perm_round(uint64_t x, const PermKey& p) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    x = (x * p.a[r] + p.b[r]) & p.mask;
    x ^= x >> p.s;
  }
  return x;
}

__host__ __device__ __forceinline__ uint64_t perm_apply(uint64_t i, const PermKey& p) {
  uint64_t x = perm_round(i, p);
  while (x >= p.n) x = perm_round(x, p);
  return x;
}

__host__ __device__ __forceinline__ uint64_t sel_threshold(double sel) {
  if (sel >= 1.0) return uint64_t(1) << 32;
  if (sel <= 0.0) return 0;
  return uint64_t(sel * 4294967296.0);
}

__host__ __device__ __forceinline__ double unit_double(uint64_t r) {
  return double(r >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace pj
