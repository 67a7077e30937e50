// pj_join.hip -- per-partition LDS hash join for gfx950.
//
//   k_join_plan  work list: partition q with |R_q| build and |S_q| probe tuples
//                gives ceil(|R_q|/base) x ceil(|S_q|/slice) items (build chunk x
//                probe slice), so heavy partitions are split on both sides.
//                Device-side, no host round trip (the reference's
//                handle_skewed_data reads the block count back to the host,
//                src/skewed_data_handler.cuh:83, and splits only the build side).
//   k_join       persistent workgroups; per item: LDS open-addressing table of
//                a power-of-two number of 16-byte key buckets with at least
//                4*|chunk| slots (occupancy <= 1/4), bucket = top bits of a
//                multiplicative hash of the key.  A build tuple claims the next
//                slot of its bucket with one LDS atomic add on the bucket's
//                claim counter, moving on to the following bucket when that one
//                is full (the reference hashes with MurmurHash3, slots with
//                int_fastdiv and inserts by CAS, src/hash_join_kernels.cuh:84-129;
//                the slot never changes the result), then the probe slice walks
//                the table (:134-181).  Slots fill each bucket in
//                order, so a probe reads one whole bucket per step (one 16-byte
//                LDS read) and stops at a match or at an empty slot: nearly
//                every probe ends in its first bucket.  Probes of one thread
//                walk interleaved, so a wave waits for its longest chain once
//                per round instead of once per probe.  Matches are
//                staged in LDS and flushed into output chunks of kOutChunk
//                pairs that each workgroup claims with one global atomic (a
//                same-address device atomic costs ~12 ns on MI355X, so claims
//                stay rare).  The empty-slot sentinel is a value absent from
//                the build chunk, so every key value can be joined (the
//                reference reserves -1, :20,86,106): in the join pipeline a
//                key whose low bits differ from the partition digit; over
//                caller-partitioned tuples one found per chunk (below the
//                chunk's minimum, above its maximum, or a gap).  A probe key
//                equal to the sentinel has no match and does not walk.
//   k_defrag_*   the last chunk of each workgroup is partly filled; pairs that
//                lie past the total are moved into those holes so the output
//                is dense in [0, total).
#include "pj_device.h"
#include "pj_internal.h"


namespace pj {

static int join_probe_per_thread(int key_bytes) { return key_bytes == 4 ? kJoinProbe4 : kJoinProbe8; }
static int join_stage_pairs(int key_bytes) { return join_probe_per_thread(key_bytes) * kJoinBlock; }
static int join_slot_bytes(int key_bytes) { return key_bytes + 4; }  // key + build payload

// Slots that fit next to the stage, a multiple of 4; for 8-byte keys at most
// 4096, so a build chunk (at most half the table) fits NB = 4 tuples per thread.
static constexpr int table_cap_of(int key_bytes) {
  const int c = ((kJoinLdsBytes - (key_bytes == 4 ? kJoinProbe4 : kJoinProbe8) * kJoinBlock * 8) / (key_bytes + 4)) & ~3;
  return key_bytes == 8 && c > 4096 ? 4096 : c;
}
static_assert(6 * kJoinBlock >= table_cap_of(4) / 2, "NB = 6 covers the largest 4-byte-key chunk");
static_assert(4 * kJoinBlock >= table_cap_of(8) / 2, "NB = 4 covers the largest 8-byte-key chunk");

int join_table_capacity(int key_bytes) { return table_cap_of(key_bytes); }

// The largest power-of-two bucket count whose slots fit table_cap: the table
// of a chunk never has more buckets.
static constexpr uint32_t cap_buckets_of(int key_bytes) {
  const uint32_t b = uint32_t(table_cap_of(key_bytes) / (16 / key_bytes));
  uint32_t p = 1;
  while (2 * p <= b) p *= 2;
  return p;
}

// LDS of a join workgroup: keys and payloads of cap_buckets buckets, one
// claim counter per bucket, then the output stage.
static size_t join_lds_bytes(int key_bytes) {
  const size_t slots = size_t(cap_buckets_of(key_bytes)) * size_t(16 / key_bytes);
  return slots * join_slot_bytes(key_bytes) + size_t(cap_buckets_of(key_bytes)) * 4 +
         size_t(join_stage_pairs(key_bytes)) * 8;
}
static_assert(2 * (4096 * 8 + 1024 * 4 + kJoinProbe4 * kJoinBlock * 8) <= 160 * 1024, "two 4-byte-key workgroups per CU");
static_assert(2 * (4096 * 12 + 2048 * 4 + kJoinProbe8 * kJoinBlock * 8) <= 160 * 1024, "two 8-byte-key workgroups per CU");

// ---------------------------------------------------------------------------
// Output staging shared by the join workgroup.
// ---------------------------------------------------------------------------
struct StageCtl {
  uint32_t n;       // pairs appended to the stage (may exceed its size: those were not written)
  uint32_t chunk;   // current output chunk
  uint32_t fill;    // pairs written to the current chunk
  uint32_t has_chunk;
  uint32_t a;       // pairs of the planned copy that go to the current chunk
  uint32_t copy_n;  // pairs of the planned copy
  uint64_t dst0, dst1;
};

// Output of the join: capacity pairs at out, then the spill area in the
// workspace for chunks claimed past the capacity (compacted afterwards).
struct OutSink {
  uint2* out;
  uint2* spill;
  uint64_t capacity;
  uint64_t spill_cap;
  uint32_t* chunk_ctr;
};

__device__ __forceinline__ void store_pair(const OutSink o, uint64_t v, uint2 pr) {
  if (v < o.capacity) o.out[v] = pr;
  else if (v - o.capacity < o.spill_cap) o.spill[v - o.capacity] = pr;
}

// Thread 0 only: plans the copy of the first min(staged, STAGE) staged pairs
// to the workgroup's output chunk, claiming a new chunk with one global atomic
// when the current one fills.  `staged` is ctl->n read after a barrier that
// follows every append; copy_stage() runs after the next barrier.
template <uint32_t STAGE>
__device__ __forceinline__ void plan_copy(const OutSink o, StageCtl* ctl, uint32_t staged) {
  const uint32_t n = staged < STAGE ? staged : STAGE;
  const uint32_t avail = ctl->has_chunk ? uint32_t(kOutChunk) - ctl->fill : 0u;
  const uint32_t a = n < avail ? n : avail;
  ctl->dst0 = uint64_t(ctl->chunk) * uint64_t(kOutChunk) + ctl->fill;
  ctl->a = a;
  ctl->copy_n = n;
  if (n > a) {
    const uint32_t c2 = atomicAdd(o.chunk_ctr, 1u);
    ctl->dst1 = uint64_t(c2) * uint64_t(kOutChunk);
    ctl->chunk = c2;
    ctl->fill = n - a;
    ctl->has_chunk = 1;
  } else {
    ctl->fill += a;
  }
}

// All threads, after the barrier that follows plan_copy(): write the planned
// pairs.  Thread 0 resets the stage count: every read of it came before that
// barrier, and appends resume only after the next one.
template <int BLOCK>
__device__ __forceinline__ void copy_stage(const OutSink o, StageCtl* ctl, const uint2* s_stage) {
  const uint32_t n = ctl->copy_n, a = ctl->a;
  const uint64_t d0 = ctl->dst0, d1 = ctl->dst1;
  if (threadIdx.x == 0) ctl->n = 0;
  if (d0 + a <= o.capacity && (n == a || d1 + (n - a) <= o.capacity)) {  // uniform: nothing goes to the spill area
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) o.out[i < a ? d0 + i : d1 + (i - a)] = s_stage[i];
    return;
  }
  for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
    const uint64_t v = i < a ? d0 + i : d1 + (i - a);
    store_pair(o, v, s_stage[i]);
  }
}

// Immediate flush: plan, barrier, copy, barrier.  Called by all threads with
// the same `staged`; ends with a barrier.
template <int BLOCK, uint32_t STAGE>
__device__ __forceinline__ void flush_stage(const OutSink o, StageCtl* ctl, const uint2* s_stage, uint32_t staged) {
  if (threadIdx.x == 0) plan_copy<STAGE>(o, ctl, staged);
  lds_barrier();
  copy_stage<BLOCK>(o, ctl, s_stage);
  lds_barrier();
}

// The keys of one 16-byte bucket.
__device__ __forceinline__ void bucket_keys(const uint4 v, uint32_t (&k)[4]) {
  k[0] = v.x; k[1] = v.y; k[2] = v.z; k[3] = v.w;
}
__device__ __forceinline__ void bucket_keys(const uint4 v, uint64_t (&k)[2]) {
  k[0] = uint64_t(v.x) | (uint64_t(v.y) << 32);
  k[1] = uint64_t(v.z) | (uint64_t(v.w) << 32);
}
__device__ __forceinline__ uint4 bucket_fill(uint32_t k) { return make_uint4(k, k, k, k); }
__device__ __forceinline__ uint4 bucket_fill(uint64_t k) {
  return make_uint4(uint32_t(k), uint32_t(k >> 32), uint32_t(k), uint32_t(k >> 32));
}

__device__ __forceinline__ uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// One work item: partition q, build chunk [rlo, rhi), probe slice [slo, shi);
// the LDS table of the chunk has 2^lg buckets (join_bucket_log2).
// Stored as two uint4 words: {q, rlo, rhi, slo}, {shi, lg, 0, 0}.
struct ItemDesc {
  uint32_t q, rlo, rhi, slo, shi, valid, lg;
};

// Bucket hash of the LDS table: a multiplicative (Fibonacci) hash whose top
// bits depend on every key bit, so keys that share their low bits (one
// partition's) still spread over the buckets.  One 32-bit multiply (two for
// 8-byte keys).
__device__ __forceinline__ uint32_t table_hash(uint32_t k) { return k * 0x9E3779B1u; }
__device__ __forceinline__ uint32_t table_hash(uint64_t k) {
  return (uint32_t(k) ^ (uint32_t(k >> 32) * 0x85EBCA77u)) * 0x9E3779B1u;
}
// First slot of key k's bucket in a table of 2^lg buckets of bs slots.
template <uint32_t BS, typename K>
__device__ __forceinline__ uint32_t first_slot(K k, uint32_t lg) {
  return BS * ((table_hash(k) >> ((32u - lg) & 31u)) & ((1u << lg) - 1u));
}

// Per-thread decode of work item `item` < total: binary search of its
// partition in item_scan, then its chunk and slice.
__device__ __forceinline__ ItemDesc decode_item(const JoinArgs& A, uint64_t item) {
  uint32_t lo = 0, hi = uint32_t(A.P);  // item_scan[lo] <= item < item_scan[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (A.item_scan[mid] <= item) lo = mid; else hi = mid;
  }
  const uint32_t q = lo;
  const uint64_t local = item - A.item_scan[q];
  const uint32_t r0 = A.roff[q], r1 = min(A.roff[q + 1], A.nr);
  const uint32_t s0 = A.soff[q], s1 = min(A.soff[q + 1], A.ns);
  const uint64_t ns = (uint64_t(s1 - s0) + A.slice - 1) / A.slice;
  const uint64_t c = local / ns, sl = local % ns;
  ItemDesc d;
  d.q = q;
  d.rlo = min(r1, r0 + uint32_t(c) * uint32_t(A.base));
  d.rhi = min(r1, d.rlo + uint32_t(A.base));
  const uint64_t slo64 = uint64_t(s0) + sl * uint64_t(A.slice);
  d.slo = uint32_t(min(uint64_t(s1), slo64));
  d.shi = uint32_t(min(uint64_t(s1), slo64 + uint64_t(A.slice)));
  d.lg = join_bucket_log2(d.rhi - d.rlo, uint32_t(A.bucket_slots), A.cap_buckets);
  d.valid = 1;
  return d;
}

// k_items_expand: descriptors of the first min(total, cap) work items, so the
// join workgroups read one descriptor per item instead of searching.
__global__ __launch_bounds__(256) void k_items_expand(JoinArgs A, uint4* desc, uint64_t cap) {
  const uint64_t total = A.item_scan[A.P];
  const uint64_t lim = total < cap ? total : cap;
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < lim; i += uint64_t(gridDim.x) * blockDim.x) {
    const ItemDesc d = decode_item(A, i);
    desc[2 * i] = make_uint4(d.q, d.rlo, d.rhi, d.slo);
    desc[2 * i + 1] = make_uint4(d.shi, d.lg, 0u, 0u);
  }
}

// An empty asm that takes a tuple's registers as read and rewritten.
__device__ __forceinline__ void tup_pin(uint2& t) { __asm__ volatile("" : "+v"(t.x), "+v"(t.y)); }
__device__ __forceinline__ void tup_pin(Tup8& t) {
  __asm__ volatile("" : "+v"(t.x), "+v"(t.y), "+v"(t.z), "+v"(t.w));
}

// Raw descriptor words of `item`, loaded without a branch (the index is
// clamped) so that nothing consumes them before desc_from_raw(): the loads
// stay in flight across the current item.
__device__ __forceinline__ void raw_desc_load(const JoinArgs& A, uint64_t item, uint4& a, uint4& b) {
  const uint64_t i = item < A.desc_cap ? item : 0;
  a = A.desc[2 * i];
  b = A.desc[2 * i + 1];
}

// Decoded descriptor, moved to scalar registers (every field is uniform).
// The raw words are read on every path (selects, not branches): a load whose
// result some path leaves unread stays "pending" for the compiler, which then
// waits for all memory operations before it reuses those registers -- at the
// top of the next item, where that wait would serialise the prefetch.
__device__ __forceinline__ ItemDesc desc_from_raw(const JoinArgs& A, const uint4& a, const uint4& b, uint64_t item,
                                                  uint64_t total) {
  const bool valid = item < total;
  ItemDesc d;
  d.q = valid ? a.x : 0u;
  d.rlo = valid ? a.y : 0u;
  d.rhi = valid ? a.z : 0u;
  d.slo = valid ? a.w : 0u;
  d.shi = valid ? b.x : 0u;
  d.lg = valid ? b.y : 0u;
  d.valid = valid ? 1u : 0u;
  if (valid && item >= A.desc_cap) d = decode_item(A, item);  // beyond the expanded descriptors
  d.q = uniform(d.q);
  d.rlo = uniform(d.rlo);
  d.rhi = uniform(d.rhi);
  d.slo = uniform(d.slo);
  d.shi = uniform(d.shi);
  d.valid = uniform(d.valid);
  d.lg = uniform(d.lg);
  return d;
}

// ---------------------------------------------------------------------------
// k_join: software-pipelined persistent join.  While item i is built and
// probed, the build tuples and first probe round of item i + grid are in
// flight into registers and the descriptor of item i + 2*grid is loading.
// Each thread holds up to NB build and R probe tuples.  A probe round runs
// steps in which every unfinished probe of the thread reads its next slot; a
// wave steps until its probes have all reached an empty slot (or, with
// UNIQUE, their match), appending matches to the LDS stage with one LDS
// atomic per wave and step.  When the stage is full the wave stops with the
// unwritten matches pending, the workgroup flushes and the round resumes.
// ---------------------------------------------------------------------------
// DIGIT_SENTINEL: every key of partition q has low bits q (the join pipeline
// partitions by the low bits itself), so q ^ 1 is absent from partition q and
// marks empty slots; otherwise (caller-partitioned tuples) each item finds a
// value absent from its build chunk.
template <int KB, int BLOCK, int NB, int R, bool UNIQUE, bool DIGIT_SENTINEL>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_join(JoinArgs A) {
  using KTr = KeyT<KB>;
  using key_t = typename KTr::key_t;
  using tup_t = typename KTr::tup_t;
  constexpr uint32_t ROUND = uint32_t(R) * BLOCK;
  constexpr uint32_t STAGE = ROUND;
  constexpr int CNT_BITS = R < 2 ? 1 : (R < 4 ? 2 : (R < 8 ? 3 : (R < 16 ? 4 : 5)));  // bits of a count <= R
  constexpr uint32_t BS = 16 / sizeof(key_t);  // slots per bucket
  static_assert(STAGE >= uint32_t(R) * kWave, "one wave's matches of a step must fit an empty stage");
  extern __shared__ __align__(16) uint8_t s_raw[];
  const uint32_t cap_slots = A.cap_buckets * BS;
  key_t* s_keys = reinterpret_cast<key_t*>(s_raw);                                        // cap_slots keys
  int32_t* s_pays = reinterpret_cast<int32_t*>(s_raw + size_t(cap_slots) * sizeof(key_t));  // their payloads
  uint32_t* s_cnt = reinterpret_cast<uint32_t*>(s_pays + cap_slots);                      // claims per bucket
  uint2* s_stage = reinterpret_cast<uint2*>(s_cnt + A.cap_buckets);                       // STAGE pairs
  __shared__ StageCtl ctl;
  __shared__ uint32_t s_any[3];
  // sentinel search (caller-partitioned tuples): chunk minimum and maximum,
  // and a bitmap of the values 1..c+1 for a chunk holding both extremes
  constexpr int kGapWords = DIGIT_SENTINEL ? 1 : (NB * BLOCK + 2 + 31) / 32;
  __shared__ key_t s_ext[2];
  __shared__ uint32_t s_gap[kGapWords + 1];
  uint32_t vote = 0;

  const int tid = threadIdx.x;
  const tup_t* rt = static_cast<const tup_t*>(A.rt);
  const tup_t* st = static_cast<const tup_t*>(A.st);
  const uint64_t total_items = A.item_scan[A.P];
  const OutSink sink{reinterpret_cast<uint2*>(A.out), reinterpret_cast<uint2*>(A.spill), uint64_t(A.capacity),
                     uint64_t(gridDim.x) * uint64_t(kOutChunk), A.chunk_ctr};
  if (tid == 0) {
    ctl.n = 0;
    ctl.fill = 0;
    ctl.chunk = 0;
    ctl.has_chunk = 0;
    ctl.copy_n = 0;
    s_any[0] = s_any[1] = s_any[2] = 0;
  }
  lds_barrier();

  uint64_t item = blockIdx.x;
  uint4 ra, rb;
  raw_desc_load(A, item, ra, rb);
  ItemDesc d = desc_from_raw(A, ra, rb, item, total_items);
  raw_desc_load(A, item + gridDim.x, ra, rb);
  ItemDesc dn = desc_from_raw(A, ra, rb, item + gridDim.x, total_items);
  tup_t bt[NB], pt[R];
  // Row k (tuples lo + k * BLOCK + tid) loads only if it starts below the
  // range's end (a uniform branch: a row wholly past it costs no memory
  // instruction) and then with indices clamped to the last tuple, without a
  // per-lane branch; lanes past the end reread that tuple and are excluded
  // by count (lanes_of) wherever used.  Rows past the end hold zeros.
  auto load_item = [&](const ItemDesc& x, tup_t (&b)[NB], tup_t (&p)[R]) {
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      if (x.rlo + uint32_t(k) * BLOCK < x.rhi) b[k] = rt[min(x.rlo + uint32_t(k) * BLOCK + tid, x.rhi - 1u)];
      else b[k] = tup_t{};
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
      if (x.slo + uint32_t(k) * BLOCK < x.shi) p[k] = st[min(x.slo + uint32_t(k) * BLOCK + tid, x.shi - 1u)];
      else p[k] = tup_t{};
    }
  };
  // Bit k set: tuple lo + k * BLOCK + tid lies below hi (k < K).
  auto lanes_of = [&](uint32_t lo, uint32_t hi, int K) -> uint32_t {
    const uint32_t n = hi > lo + uint32_t(tid) ? (hi - lo - uint32_t(tid) + BLOCK - 1u) / BLOCK : 0u;
    return n >= uint32_t(K) ? (1u << K) - 1u : (1u << n) - 1u;
  };
  if (d.valid) load_item(d, bt, pt);
  // A value absent from the item's build chunk, for caller-partitioned
  // tuples: the chunk's minimum - 1, else its maximum + 1, else (the chunk
  // holds both extremes) the smallest of 1..c+1 it lacks, which exists since
  // the chunk has c keys.  Uniform result; every branch is uniform.
  auto chunk_sentinel = [&](const ItemDesc& x) -> key_t {
    using ull = unsigned long long;
    key_t mn = ~key_t(0), mx = 0;
#pragma unroll
    for (int k = 0; k < NB; ++k)
      if (x.rlo + uint32_t(k) * BLOCK + tid < x.rhi) {
        const key_t v = KTr::key(bt[k]);
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
      }
    for (int off = kWave / 2; off > 0; off >>= 1) {
      const key_t a = key_t(__shfl_xor(mn, off, kWave)), b = key_t(__shfl_xor(mx, off, kWave));
      mn = a < mn ? a : mn;
      mx = b > mx ? b : mx;
    }
    if (tid == 0) {
      s_ext[0] = ~key_t(0);
      s_ext[1] = 0;
    }
    lds_barrier();
    if (lane_id() == 0) {
      if constexpr (sizeof(key_t) == 8) {
        atomicMin(reinterpret_cast<ull*>(&s_ext[0]), ull(mn));
        atomicMax(reinterpret_cast<ull*>(&s_ext[1]), ull(mx));
      } else {
        atomicMin(&s_ext[0], mn);
        atomicMax(&s_ext[1], mx);
      }
    }
    lds_barrier();
    const key_t gmn = s_ext[0], gmx = s_ext[1];
    if (gmn != key_t(0)) return gmn - 1;
    if (gmx != ~key_t(0)) return gmx + 1;
    const uint32_t c = x.rhi - x.rlo, words = (c + 2u + 31u) / 32u;  // bits 0..c+1
    for (uint32_t i = tid; i < words; i += BLOCK) s_gap[i] = 0;
    if (tid == 0) s_gap[kGapWords] = 0xFFFFFFFFu;
    lds_barrier();
#pragma unroll
    for (int k = 0; k < NB; ++k)
      if (x.rlo + uint32_t(k) * BLOCK + tid < x.rhi) {
        const key_t v = KTr::key(bt[k]);
        if (v >= key_t(1) && v <= key_t(c + 1u)) atomicOr(&s_gap[uint32_t(v) / 32u], 1u << (uint32_t(v) % 32u));
      }
    lds_barrier();
    for (uint32_t i = tid; i < words; i += BLOCK) {
      uint32_t free_bits = ~s_gap[i];
      if (i == 0) free_bits &= ~1u;                                   // 0 is present
      const uint32_t top = c + 1u - i * 32u;                          // highest candidate bit in this word
      if (top < 31u) free_bits &= (2u << top) - 1u;
      if (free_bits) atomicMin(&s_gap[kGapWords], i * 32u + uint32_t(__builtin_ctz(free_bits)));
    }
    lds_barrier();
    return key_t(s_gap[kGapWords]);
  };

  while (d.valid) {
    // The empty asm redefines this item's tuples (copied from the previous
    // prefetch) before the next prefetch is issued: the compiler's wait for
    // them then falls here, and not in the build, where it would also wait
    // for the next item's loads.
#pragma unroll
    for (int k = 0; k < NB; ++k) tup_pin(bt[k]);
#pragma unroll
    for (int k = 0; k < R; ++k) tup_pin(pt[k]);
    const uint32_t tsize = BS << d.lg, tmask = tsize - 1u;
    auto fill_table = [&](key_t empty) {
      const uint4 fill = bucket_fill(empty);
      uint4* s4 = reinterpret_cast<uint4*>(s_keys);
      for (uint32_t i = tid; i < tsize / BS; i += BLOCK) {
        s4[i] = fill;
        s_cnt[i] = 0u;
      }
    };
    key_t sentinel = key_t(d.q ^ 1u);
    tup_t bn[NB], pn[R];
    const uint64_t item_nn = item + 2 * uint64_t(gridDim.x);
    if constexpr (DIGIT_SENTINEL) {
      // The table is filled before the next prefetch is issued: a register
      // reload at the fill (the multi-match kernels run out of registers)
      // then waits for nothing that is still in flight.
      fill_table(sentinel);
      if (dn.valid) load_item(dn, bn, pn);
      raw_desc_load(A, item_nn, ra, rb);  // decoded at the end of this item
    } else {
      if (dn.valid) load_item(dn, bn, pn);
      raw_desc_load(A, item_nn, ra, rb);
      sentinel = chunk_sentinel(d);
      fill_table(sentinel);
    }
    lds_barrier();
    // The previous item's matches go out now, after this item's prefetch was
    // issued: the in-order memory counter that the next wait for the prefetch
    // uses then finds these stores long complete.
    copy_stage<BLOCK>(sink, &ctl, s_stage);
    // Build: each tuple takes the next slot of its bucket from the bucket's
    // claim counter (one LDS atomic add), moving on to the following bucket
    // when its claim is past the bucket's end.  The claims of one bucket are
    // consecutive, so its filled slots are a prefix of it, and every bucket a
    // key passed over is full: a probe that walks from the key's bucket to
    // the first bucket with an empty slot meets every copy of the key.
    bool dup = false;
    {
      const uint32_t bmask = (1u << d.lg) - 1u;
      uint32_t bkt[NB];
      uint32_t pend = lanes_of(d.rlo, d.rhi, NB);
#pragma unroll
      for (int k = 0; k < NB; ++k) bkt[k] = first_slot<BS>(KTr::key(bt[k]), d.lg) / BS;
      const uint32_t valid = pend;
      while (pend) {
        uint32_t pos[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k)
          if (pend & (1u << k)) pos[k] = atomicAdd(&s_cnt[bkt[k]], 1u);
#pragma unroll
        for (int k = 0; k < NB; ++k)
          if (pend & (1u << k)) {
            if (pos[k] < BS) {
              const uint32_t sl = bkt[k] * BS + pos[k];
              s_keys[sl] = KTr::key(bt[k]);
              s_pays[sl] = KTr::pay(bt[k]);
              pend &= ~(1u << k);
            } else {
              bkt[k] = (bkt[k] + 1u) & bmask;
            }
          }
      }
      if (!UNIQUE) {
        // Duplicate keys in the chunk: each tuple counts its key over the
        // buckets from its own bucket to the one it landed in.  Equal keys
        // share their first bucket, so the copy that landed furthest along
        // counts every copy, itself included.  The home buckets are read
        // together without branches (a tuple's home bucket is recomputed from
        // its key rather than kept in a register); the few tuples that landed
        // further on walk on one bucket at a time.
        lds_barrier();
        constexpr int G = 3;
        uint32_t far = 0;
#pragma unroll
        for (int g = 0; g < NB; g += G) {
          uint4 hq[G];
#pragma unroll
          for (int j = 0; j < G; ++j)
            if (g + j < NB) hq[j] = *reinterpret_cast<const uint4*>(&s_keys[first_slot<BS>(KTr::key(bt[g + j]), d.lg)]);
#pragma unroll
          for (int j = 0; j < G; ++j)
            if (g + j < NB) {
              const int k = g + j;
              key_t kk[BS];
              bucket_keys(hq[j], kk);
              const key_t key = KTr::key(bt[k]);
              uint32_t n = 0;
#pragma unroll
              for (int i = 0; i < int(BS); ++i) n += kk[i] == key ? 1u : 0u;
              const bool v = (valid >> k) & 1u;
              dup |= v && n > 1u;
              far |= v && bkt[k] != first_slot<BS>(key, d.lg) / BS ? (1u << k) : 0u;
            }
        }
        if (far) {
#pragma unroll
          for (int k = 0; k < NB; ++k)
            if (far & (1u << k)) {
              const key_t key = KTr::key(bt[k]);
              uint32_t b = first_slot<BS>(key, d.lg) / BS, n = 0;
              while (true) {
                key_t kk[BS];
                bucket_keys(*reinterpret_cast<const uint4*>(&s_keys[b * BS]), kk);
#pragma unroll
                for (int i = 0; i < int(BS); ++i) n += kk[i] == key ? 1u : 0u;
                if (b == bkt[k]) break;
                b = (b + 1u) & bmask;
              }
              dup |= n > 1u;
            }
        }
      }
    }
    // Without duplicate keys in the chunk a probe has at most one match.  The
    // vote is also the barrier between the build and the probes, so it runs
    // whether or not the plan already promises unique keys.
    const bool chunk_has_dup = block_any(dup, s_any, vote++);
    const bool first_match_only = UNIQUE || !chunk_has_dup;

    // A probe round of R tuples per thread with at most one match per probe:
    // each step reads one bucket per unfinished probe, which ends at a match
    // or at an empty slot; then all matches are appended at once (the stage
    // is empty and holds R per thread).  Ends with a barrier.
    auto round_single = [&](const tup_t (&ptr)[R], uint32_t r0) {
      uint32_t slot[R];
      uint32_t walk = 0, found = 0;
      walk = lanes_of(r0, d.shi, R);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        if constexpr (!DIGIT_SENTINEL) walk &= KTr::key(ptr[k]) != sentinel ? ~0u : ~(1u << k);
        slot[k] = first_slot<BS>(KTr::key(ptr[k]), d.lg);
      }
      // The first step, taken by nearly every probe, without branches: the
      // bucket reads go out in groups of G (an invalid probe reads bucket 0
      // and is ignored) and are examined with selects.  Later steps, taken by
      // few probes, branch per probe so that finished probes read nothing.
      {
        constexpr int G = 3;
#pragma unroll
        for (int g = 0; g < R; g += G) {
          uint4 bq[G];
#pragma unroll
          for (int j = 0; j < G; ++j)
            if (g + j < R) bq[j] = *reinterpret_cast<const uint4*>(&s_keys[slot[g + j]]);
#pragma unroll
          for (int j = 0; j < G; ++j)
            if (g + j < R) {
              const int k = g + j;
              key_t kk[BS];
              bucket_keys(bq[j], kk);
              const key_t key = KTr::key(ptr[k]);
              const bool active = ((walk >> k) & 1u) != 0u;
              const bool empty = kk[BS - 1] == sentinel;
              uint32_t hit = BS;
#pragma unroll
              for (int i = BS - 1; i >= 0; --i) hit = kk[i] == key ? uint32_t(i) : hit;
              const bool is_hit = active && hit < BS;
              const bool cont = active && hit >= BS && !empty;
              found |= is_hit ? (1u << k) : 0u;
              walk &= cont ? ~0u : ~(1u << k);
              slot[k] = is_hit ? slot[k] + hit : (cont ? ((slot[k] + BS) & tmask) : slot[k]);
            }
        }
      }
      while (__ballot(walk != 0)) {
#pragma unroll
        for (int k = 0; k < R; ++k)
          if (walk & (1u << k)) {
            key_t kk[BS];
            bucket_keys(*reinterpret_cast<const uint4*>(&s_keys[slot[k]]), kk);
            const key_t key = KTr::key(ptr[k]);
            // A bucket's filled slots are a prefix of it: keys start at a
            // bucket's first slot and only move on past occupied slots, and
            // the table (a whole number of buckets) wraps to bucket 0.  So
            // the bucket has an empty slot iff its last slot is empty.
            const bool empty = kk[BS - 1] == sentinel;
            uint32_t hit = BS;
#pragma unroll
            for (int i = BS - 1; i >= 0; --i) hit = kk[i] == key ? uint32_t(i) : hit;
            if (hit < BS) {
              found |= 1u << k;
              walk &= ~(1u << k);
              slot[k] += hit;
            } else if (empty) {
              walk &= ~(1u << k);
            } else {
              slot[k] = (slot[k] + BS) & tmask;
            }
          }
      }
      if (__ballot(found != 0)) {
        const uint32_t cnt = __popc(found);
        uint32_t excl = 0, wtotal = 0;
#pragma unroll
        for (int b = 0; b < CNT_BITS; ++b) {
          const uint64_t m = __ballot((cnt >> b) & 1u);
          excl += mbcnt(m) << b;
          wtotal += uint32_t(__popcll(m)) << b;
        }
        uint32_t wbase = 0;
        if (lane_id() == 0) wbase = atomicAdd(&ctl.n, wtotal);
        uint32_t pos = uniform(wbase) + excl;
#pragma unroll
        for (int k = 0; k < R; ++k)
          if ((found & (1u << k)) && pos < STAGE)
            s_stage[pos++] = make_uint2(uint32_t(s_pays[slot[k]]), uint32_t(KTr::pay(ptr[k])));
      }
      lds_barrier();
    };

    // A probe round with any number of matches per probe: steps advance every
    // unfinished probe to its next match or an empty slot, appending matches
    // with one LDS atomic per wave and step.  When the stage is full the wave
    // stops with the unwritten matches pending, the workgroup flushes and the
    // round resumes.  Ends with a barrier; the last matches stay staged.
    auto round_multi = [&](const tup_t (&ptr)[R], uint32_t r0) {
      uint32_t slot[R];
      uint32_t walk = 0;  // bit k: probe k has not reached its end
      walk = lanes_of(r0, d.shi, R);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        if constexpr (!DIGIT_SENTINEL) walk &= KTr::key(ptr[k]) != sentinel ? ~0u : ~(1u << k);
        slot[k] = first_slot<BS>(KTr::key(ptr[k]), d.lg);
      }
      while (true) {
        while (__ballot(walk != 0)) {
          key_t e[R];
#pragma unroll
          for (int k = 0; k < R; ++k)
            if (walk & (1u << k)) e[k] = s_keys[slot[k]];
          uint32_t hit = 0;
#pragma unroll
          for (int k = 0; k < R; ++k)
            if (walk & (1u << k)) {
              if (e[k] == KTr::key(ptr[k])) hit |= 1u << k;
              else if (e[k] == sentinel) walk &= ~(1u << k);
              else slot[k] = (slot[k] + 1u) & tmask;
            }
          if (__ballot(hit != 0)) {
            // wave prefix sum of the per-lane match counts from CNT_BITS ballots
            const uint32_t cnt = __popc(hit);
            uint32_t excl = 0, wtotal = 0;
#pragma unroll
            for (int b = 0; b < CNT_BITS; ++b) {
              const uint64_t m = __ballot((cnt >> b) & 1u);
              excl += mbcnt(m) << b;
              wtotal += uint32_t(__popcll(m)) << b;
            }
            uint32_t wbase = 0;
            if (lane_id() == 0) wbase = atomicAdd(&ctl.n, wtotal);
            wbase = uniform(wbase);
            uint32_t pos = wbase + excl;
#pragma unroll
            for (int k = 0; k < R; ++k)
              if (hit & (1u << k)) {
                if (pos < STAGE) {
                  s_stage[pos] = make_uint2(uint32_t(s_pays[slot[k]]), uint32_t(KTr::pay(ptr[k])));
                  slot[k] = (slot[k] + 1u) & tmask;
                }
                ++pos;  // a match past the stage stays pending at its slot
              }
            if (wbase + wtotal > STAGE) break;  // stage full: flush, then resume
          }
        }
        // A wave leaves its steps with probes pending only after filling the
        // stage, so `any` implies staged > 0.
        const bool any = block_any(walk != 0, s_any, vote++);
        if (!any) break;
        flush_stage<BLOCK, STAGE>(sink, &ctl, s_stage, ctl.n);
      }
    };

    auto round = [&](const tup_t (&ptr)[R], uint32_t r0) {
      if (first_match_only) round_single(ptr, r0);
      else round_multi(ptr, r0);
    };
    round(pt, d.slo);
    for (uint32_t r0 = d.slo + ROUND; r0 < d.shi; r0 += ROUND) {
      // another round of this item: empty the stage first
      flush_stage<BLOCK, STAGE>(sink, &ctl, s_stage, ctl.n);
      tup_t pr[R];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        if (r0 + uint32_t(k) * BLOCK < d.shi) pr[k] = st[min(r0 + uint32_t(k) * BLOCK + tid, d.shi - 1u)];
        else pr[k] = tup_t{};
      }
      round(pr, r0);
    }
    // The round's closing barrier follows every table read and stage append:
    // the next item may overwrite the table, and the stage is copied out after
    // the next item's first barrier.
    if (tid == 0) plan_copy<STAGE>(sink, &ctl, ctl.n);
    // advance the pipeline.  The empty asm keeps the raw descriptor words in
    // VGPRs until here: the compiler would otherwise move them to scalar
    // registers right after their loads, which waits for every load issued
    // before them -- the next item's tuples included.
    __asm__ volatile("" : "+v"(ra.x), "+v"(ra.y), "+v"(ra.z), "+v"(ra.w), "+v"(rb.x), "+v"(rb.y), "+v"(rb.z), "+v"(rb.w));
    d = dn;
    dn = desc_from_raw(A, ra, rb, item_nn, total_items);
#pragma unroll
    for (int k = 0; k < NB; ++k) bt[k] = bn[k];
#pragma unroll
    for (int k = 0; k < R; ++k) pt[k] = pn[k];
    item += gridDim.x;
  }
  lds_barrier();
  copy_stage<BLOCK>(sink, &ctl, s_stage);
  if (tid == 0)
    A.wg_last[blockIdx.x] = ctl.has_chunk ? ((uint64_t(ctl.chunk) << 32) | ctl.fill) : ~0ull;
}

// ---------------------------------------------------------------------------
// Defragmentation.  Chunks 0..C-1 were claimed; every chunk but the last one of
// each workgroup is full.  total = (C - W) * CH + sum(fill of last chunks).
// Holes: [c*CH + fill, (c+1)*CH) of each last chunk, below total.  Strays:
// pairs at positions >= total (the complement of the holes in [total, C*CH)).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_defrag_plan(const uint32_t* chunk_ctr, const uint64_t* wg_last, int nwg,
                                                      int64_t capacity, int64_t* d_count, DefragState* st,
                                                      uint64_t* hole_lo, uint64_t* hole_pre, uint64_t* stray_lo,
                                                      uint64_t* stray_pre) {
  __shared__ uint64_t s_rec[1024];
  __shared__ uint64_t s_sorted[1024];
  __shared__ uint64_t s_tmp[1024 / kWave + 1];
  const int tid = threadIdx.x;
  const uint64_t CH = uint64_t(kOutChunk);
  const uint64_t rec = tid < nwg ? wg_last[tid] : ~0ull;
  const bool valid = rec != ~0ull;
  s_rec[tid] = rec;
  __syncthreads();
  if (valid) {  // rank by chunk id (chunk ids are unique)
    uint32_t rank = 0;
    for (int w = 0; w < nwg; ++w) rank += (s_rec[w] != ~0ull && s_rec[w] < rec) ? 1u : 0u;
    s_sorted[rank] = rec;
  }
  uint64_t fill_sum;
  block_exclusive_scan64<1024>(valid ? (rec & 0xffffffffull) : 0ull, s_tmp, &fill_sum);
  const uint32_t W = uint32_t(__syncthreads_count(valid));
  const uint64_t C = *chunk_ctr;
  const uint64_t total = (C - W) * CH + fill_sum;
  if (tid == 0) {
    *d_count = int64_t(total);
    st->total = total;
  }
  if (total > uint64_t(capacity)) {
    if (tid == 0) {
      st->moves = 0;
      st->n_holes = 0;
      st->n_strays = 0;
      hole_pre[0] = 0;
      stray_pre[0] = 0;
    }
    return;
  }
  // record i (sorted): data [c*CH, c*CH + f), hole [c*CH + f, (c+1)*CH)
  uint64_t hlen = 0, hlo = 0, slen = 0, slo = 0;
  const uint32_t i = uint32_t(tid);
  if (i < W) {
    const uint64_t c = s_sorted[i] >> 32, f = s_sorted[i] & 0xffffffffull;
    const uint64_t plo = c * CH + f, phi = (c + 1) * CH;
    hlo = plo;
    hlen = plo < total ? min(phi, total) - plo : 0;
    const uint64_t prev_phi = i == 0 ? 0 : ((s_sorted[i - 1] >> 32) + 1) * CH;
    slo = prev_phi > total ? prev_phi : total;  // data gap before this record's hole, above total
    slen = plo > slo ? plo - slo : 0;
  } else if (i == W) {  // tail above the last hole
    const uint64_t last_phi = W == 0 ? 0 : ((s_sorted[W - 1] >> 32) + 1) * CH;
    slo = last_phi > total ? last_phi : total;
    slen = C * CH > slo ? C * CH - slo : 0;
  }
  __syncthreads();
  uint64_t hsum, ssum;
  const uint64_t hpre = block_exclusive_scan64<1024>(hlen, s_tmp, &hsum);
  __syncthreads();
  const uint64_t spre = block_exclusive_scan64<1024>(slen, s_tmp, &ssum);
  if (i < W) {
    hole_lo[i] = hlo;
    hole_pre[i] = hpre;
  }
  if (i <= W) {
    stray_lo[i] = slo;
    stray_pre[i] = spre;
  }
  if (tid == 0) {
    hole_pre[W] = hsum;
    stray_pre[W + 1] = ssum;
    st->n_holes = W;
    st->n_strays = W + 1;
    st->moves = hsum < ssum ? hsum : ssum;
  }
}

__device__ __forceinline__ uint32_t upper_index(const uint64_t* pre, uint32_t n, uint64_t k) {
  // largest i < n with pre[i] <= k
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pre[mid] <= k) lo = mid; else hi = mid;
  }
  return lo;
}

// Workgroup i moves stray run i: moves k in [stray_pre[i], stray_pre[i+1])
// (below `moves`) go from stray_lo[i] + (k - stray_pre[i]) to the hole holding
// move k.  One search finds the run's first hole; the workgroup then walks the
// holes in order, copying each overlap with contiguous reads and writes.
// Every index is uniform in the workgroup; each thread reads four pairs
// before it writes them, so a run's copy keeps many loads in flight, and a
// long run is split between several workgroups.
__global__ __launch_bounds__(256) void k_defrag_move(const DefragState* st, const uint64_t* hole_lo,
                                                     const uint64_t* hole_pre, const uint64_t* stray_lo,
                                                     const uint64_t* stray_pre, uint2* out, int64_t capacity,
                                                     const uint2* spill) {
  const uint64_t moves = st->moves;
  const uint32_t nh = st->n_holes, ns = st->n_strays;
  const uint32_t i = blockIdx.x;
  if (i >= ns || moves == 0) return;
  const uint64_t r0 = stray_pre[i], r1 = min(stray_pre[i + 1], moves);
  if (r0 >= r1) return;
  const uint64_t s0 = stray_lo[i] - r0;  // source of move k: s0 + k
  // blockIdx.y takes one of gridDim.y equal slices of the run
  const uint64_t sl = (r1 - r0 + gridDim.y - 1) / gridDim.y;
  uint64_t k = r0 + sl * blockIdx.y;
  const uint64_t kend = min(r1, k + sl);
  if (k >= kend) return;
  // k < moves <= hole_pre[nh], so the hole of every move here is below nh
  uint32_t h = upper_index(hole_pre, nh, k);
  while (k < kend) {
    const uint64_t hend = hole_pre[h + 1];
    const uint64_t d0 = hole_lo[h] - hole_pre[h];  // destination of move k: d0 + k
    const uint64_t e = min(kend, hend);
    constexpr uint32_t U = 4;
    for (uint64_t base = k; base < e; base += uint64_t(U) * blockDim.x) {
      uint2 v[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint64_t kk = base + threadIdx.x + uint64_t(u) * blockDim.x;
        const uint64_t src = s0 + kk;
        if (kk < e) v[u] = src < uint64_t(capacity) ? out[src] : spill[src - uint64_t(capacity)];
      }
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint64_t kk = base + threadIdx.x + uint64_t(u) * blockDim.x;
        if (kk < e) out[d0 + kk] = v[u];
      }
    }
    k = e;
    ++h;
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int KB, int NB, int R, bool DS>
static void join_launch_t(const JoinArgs& a, int unique, size_t lds, hipStream_t s) {
  dim3 grid(kJoinGrid), block(kJoinBlock);
  if (unique) hipLaunchKernelGGL((k_join<KB, kJoinBlock, NB, R, true, DS>), grid, block, lds, s, a);
  else hipLaunchKernelGGL((k_join<KB, kJoinBlock, NB, R, false, DS>), grid, block, lds, s, a);
}

template <bool DS>
static void launch_join_kernel_t(const JoinArgs& A, int key_bytes, int base, int unique, size_t lds, hipStream_t s) {
  // build tuples per thread: NB * kJoinBlock >= base (base <= table_cap / 2 - 1)
  if (key_bytes == 4) {
    if (base <= 3 * kJoinBlock) join_launch_t<4, 3, kJoinProbe4, DS>(A, unique, lds, s);
    else if (base <= 5 * kJoinBlock) join_launch_t<4, 5, kJoinProbe4, DS>(A, unique, lds, s);
    else join_launch_t<4, 6, kJoinProbe4, DS>(A, unique, lds, s);
  } else {
    if (base <= 2 * kJoinBlock) join_launch_t<8, 2, kJoinProbe8, DS>(A, unique, lds, s);
    else join_launch_t<8, 4, kJoinProbe8, DS>(A, unique, lds, s);
  }
}

// digit_sentinel: the tuples were partitioned by their low key bits here, so
// partition q's keys all have low bits q (see k_join).
static hipError_t launch_join_kernel(const JoinArgs& A, int key_bytes, int base, int unique, bool digit_sentinel,
                                     hipStream_t s) {
  const size_t lds = join_lds_bytes(key_bytes);
  if (digit_sentinel) launch_join_kernel_t<true>(A, key_bytes, base, unique, lds, s);
  else launch_join_kernel_t<false>(A, key_bytes, base, unique, lds, s);
  return hipGetLastError();
}

static JoinArgs join_args(const Layout& L, int key_bytes, int base, int64_t slice, pj_pair* out, int64_t capacity) {
  JoinArgs A{};
  const int last = L.pp.passes == 2;
  A.rt = last ? L.t[0].buf2 : L.t[0].buf1;
  A.st = last ? L.t[1].buf2 : L.t[1].buf1;
  A.roff = L.t[0].off;
  A.soff = L.t[1].off;
  A.P = L.pp.P;
  A.base = base;
  A.slice = slice;
  A.item_scan = L.item_scan;
  A.out = out;
  A.capacity = capacity;
  A.spill = L.spill;
  A.chunk_ctr = L.chunk_ctr;
  A.wg_last = L.wg_last;
  A.nr = L.t[0].n;
  A.ns = L.t[1].n;
  A.desc = reinterpret_cast<const uint4*>(L.desc);
  A.desc_cap = L.desc_cap;
  A.bucket_slots = 16 / key_bytes;
  A.cap_buckets = cap_buckets_of(key_bytes);
  return A;
}

hipError_t launch_join(const Layout& L, int key_bytes, int base, int64_t slice, int unique, pj_pair* out,
                       int64_t capacity, hipStream_t s, bool digit_sentinel) {
  JoinArgs A = join_args(L, key_bytes, base, slice, out, capacity);
  hipLaunchKernelGGL(k_items_expand, dim3(1024), dim3(256), 0, s, A, reinterpret_cast<uint4*>(L.desc), L.desc_cap);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_join_kernel(A, key_bytes, base, unique, digit_sentinel, s);
}

// Workgroups per stray run in the output defrag, each copying one slice of
// it: the copy is latency-bound (a search, then a walk over the holes), so
// more, shorter slices finish sooner (compact 0.065 -> 0.037 ms at 64M x 256M
// going from 4 to 16).
constexpr int kDefragSplit = 16;

hipError_t launch_defrag(const Layout& L, pj_pair* out, int64_t capacity, int64_t* d_count, hipStream_t s) {
  hipLaunchKernelGGL(k_defrag_plan, dim3(1), dim3(1024), 0, s, L.chunk_ctr, L.wg_last, kJoinGrid, capacity, d_count,
                     L.dstate, L.hole_lo, L.hole_pre, L.stray_lo, L.stray_pre);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_defrag_move, dim3(kJoinGrid + 1, kDefragSplit), dim3(256), 0, s, L.dstate, L.hole_lo, L.hole_pre, L.stray_lo,
                     L.stray_pre, reinterpret_cast<uint2*>(out), capacity, reinterpret_cast<const uint2*>(L.spill));
  return hipGetLastError();
}

// Join over externally partitioned tuples (phase API): workspace carve-up.
size_t join_generic_ws_bytes(int B) {
  const size_t P = size_t(1) << B;
  size_t b = 0;
  b += (P + 1) * 8;                                   // item_scan
  b += 256;                                           // chunk counter + defrag state
  b += size_t(kJoinGrid) * 8;                         // wg_last
  b += 4 * (size_t(kJoinGrid) + 3) * 8;               // hole/stray arrays
  b += size_t(kJoinGrid) * size_t(kOutChunk) * 8;     // spill
  b += 2 * 4096 * 8;                                  // scan block totals
  b += (2 * P + 65536) * 32;                          // work item descriptors
  return b + 4096;
}

hipError_t launch_join_generic(const void* rt, const uint32_t* roff, const void* st, const uint32_t* soff,
                               int key_bytes, int B, int base, int64_t slice, int unique, pj_pair* out,
                               int64_t capacity, int64_t* d_count, char* ws, hipStream_t s) {
  const size_t P = size_t(1) << B;
  auto take = [&](size_t bytes) {
    char* p = ws;
    ws += (bytes + 255) & ~size_t(255);
    return p;
  };
  Layout L{};
  L.pp.P = int(P);
  L.pp.passes = 1;
  L.t[0].off = const_cast<uint32_t*>(roff);
  L.t[1].off = const_cast<uint32_t*>(soff);
  L.t[0].buf1 = const_cast<void*>(rt);
  L.t[1].buf1 = const_cast<void*>(st);
  L.t[0].n = 0xFFFFFFFFu;  // ranges come from the caller's offsets
  L.t[1].n = 0xFFFFFFFFu;
  L.item_scan = reinterpret_cast<uint64_t*>(take((P + 1) * 8));
  L.chunk_ctr = reinterpret_cast<uint32_t*>(take(64));
  L.dstate = reinterpret_cast<DefragState*>(take(sizeof(DefragState)));
  L.wg_last = reinterpret_cast<uint64_t*>(take(size_t(kJoinGrid) * 8));
  L.hole_lo = reinterpret_cast<uint64_t*>(take((size_t(kJoinGrid) + 3) * 8));
  L.hole_pre = reinterpret_cast<uint64_t*>(take((size_t(kJoinGrid) + 3) * 8));
  L.stray_lo = reinterpret_cast<uint64_t*>(take((size_t(kJoinGrid) + 3) * 8));
  L.stray_pre = reinterpret_cast<uint64_t*>(take((size_t(kJoinGrid) + 3) * 8));
  L.spill = reinterpret_cast<pj_pair*>(take(size_t(kJoinGrid) * size_t(kOutChunk) * 8));
  L.scan_scratch = take(2 * 4096 * 8);
  L.desc_cap = 2 * P + 65536;
  L.desc = take(L.desc_cap * 32);
  hipError_t e = launch_join_plan(L, base, slice, s);
  if (e != hipSuccess) return e;
  e = launch_join(L, key_bytes, base, slice, unique, out, capacity, s, false);
  if (e != hipSuccess) return e;
  return launch_defrag(L, out, capacity, d_count, s);
}

}  // namespace pj
