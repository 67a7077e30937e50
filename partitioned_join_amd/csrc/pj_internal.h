// pj_internal.h -- structures shared by the kernel files and the host orchestration.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/pj.h"

namespace pj {

// Tuning constants (gfx950: 256 CUs, 160 KB LDS per CU, wave64).
constexpr int kHistBlock = 1024;      // combined histogram workgroup
constexpr int kScatterBlock = 1024;   // partition pass workgroup (one per CU: 16 waves)
constexpr int kScatterTile4 = 16384;  // largest scatter tile, 4-byte keys (128 KB of LDS; scatter_cfg picks the tile per fan-out)
constexpr int kScatterTile8 = 8192;   // 8-byte keys (128 KB)
constexpr int kHeavyLanes = 8;        // lanes of a wave sharing one digit or bin from which ranks and counts are aggregated
constexpr int kJoinBlock = 512;       // join workgroup
constexpr int kJoinProbe4 = 9;        // probe tuples per thread and round, 4-byte keys (4608 per round)
constexpr int kJoinProbe8 = 4;        // 8-byte keys (2048 per round)
constexpr int kJoinLdsBytes = 81408;  // dynamic LDS per join workgroup: two workgroups share a CU's 160 KB
constexpr int64_t kOutChunk = 16384;  // output pairs claimed per global allocation
constexpr int kJoinGrid = 512;        // persistent join workgroups (2 per CU)
constexpr int kPass2WgPerXcd = 16;    // pass-2 persistent workgroups per XCD and table (one per CU)
constexpr int kPass2ItemTiles = 4;    // pass-2 items hold about this many full tiles
constexpr int kMaxCombinedBits = 16;  // combined histogram limit (65536 partitions in LDS)
constexpr int kMaxTotalBits = 24;     // two passes of at most 12 bits; above kMaxCombinedBits the
                                      // second pass counts its digits over pass 1's output ("deep" plans)
constexpr int kMaxRadix = 4096;       // largest fan-out of one pass (12 bits)
constexpr int kMaxSegPerXcd = 32;    // segments per XCD and table at most: the [segment][digit] matrix the histogram writes and the column scans read shrinks with it, while pass 1 (one workgroup per segment) still has two per CU (64M x 256M: 64 -> 4.41 ms, 32 -> 4.17, 16 -> 4.39)

// One table being partitioned.  All pointers are device pointers.
struct PartTable {
  const void* keys;      // SoA keys (pass-1 input)
  const int32_t* pay;    // SoA payload or nullptr (row ids)
  uint32_t n;            // rows
  uint32_t nseg;         // histogram / pass-1 segments (multiple of 8)
  void* buf1;            // pass-1 output tuples
  void* buf2;            // pass-2 output tuples (2-pass plans)
  uint32_t* mat;         // [nseg][P] per-segment fine counts; exclusive column scan after colscan
  uint32_t* matc;        // [nseg][F1] per-segment pass-1 counts (2-pass); scanned likewise
  uint32_t* tot;         // [P] fine totals
  uint32_t* totc;        // [F1] pass-1 totals (2-pass)
  uint32_t* off;         // [P+1] final partition starts, off[P] = n
  // deep plans (total bits > kMaxCombinedBits) only:
  uint32_t* offc;        // [F1+1] pass-1 partition starts
  uint32_t* h2;          // [items][F2] pass-2 digit counts per pass-2 item, then their exclusive scan
                         // over the items of the same pass-1 partition
  uint32_t* ipre;        // [RN+1] pass-2 items before each pass-1 partition rank
};

struct PartParams {
  int B;        // total radix bits
  int b1, b2;   // pass bits (b2 == 0 for one pass)
  int P, F1, F2;
  int passes;
  int shift0;   // every digit is taken from key >> shift0 (0 in the join pipeline)
};

// Segment s of nseg splits [0, n) at multiples of 64 rows.
__host__ __device__ inline uint32_t seg_bound(uint32_t n, uint32_t nseg, uint32_t k) {
  if (k == 0) return 0;
  if (k >= nseg) return n;
  return uint32_t((uint64_t(n) * k / nseg) & ~uint64_t(63));
}

struct JoinArgs {
  const void* rt;          // build tuples (partitioned)
  const uint32_t* roff;    // [P+1]
  const void* st;          // probe tuples (partitioned)
  const uint32_t* soff;    // [P+1]
  int P;
  int base;                // build chunk per item
  int64_t slice;           // probe slice per item
  const uint64_t* item_scan;  // [P+1]
  pj_pair* out;
  int64_t capacity;
  pj_pair* spill;          // kJoinGrid * kOutChunk pairs
  uint32_t* chunk_ctr;     // claimed output chunks
  uint64_t* wg_last;       // [grid] (chunk << 32 | fill) or ~0 when no chunk; fill <= kOutChunk
  int bucket_slots;        // keys per 16-byte bucket: 4 (4-byte keys) or 2 (8-byte keys)
  uint32_t cap_buckets;    // largest power-of-two bucket count within join_table_capacity() slots
  uint32_t nr, ns;         // table sizes (bounds for the item ranges)
  const uint4* desc;       // expanded work item descriptors (2 uint4 per item)
  uint64_t desc_cap;       // items with a descriptor; later items are decoded by search
};

struct DefragState {
  uint64_t total;          // pairs found
  uint64_t moves;          // pairs to move
  uint32_t n_holes, n_strays;
  uint32_t pad[2];
};

// Workspace carve-up for one join.
struct Layout {
  PartTable t[2];
  int ntables;         // 1 or 2 (phase API runs one table)
  PartParams pp;
  uint64_t* item_scan;
  uint32_t* chunk_ctr;
  uint64_t* wg_last;
  pj_pair* spill;
  DefragState* dstate;
  uint64_t* hole_lo;   // [kJoinGrid + 2]
  uint64_t* hole_pre;  // [kJoinGrid + 3]
  uint64_t* stray_lo;  // [kJoinGrid + 2]
  uint64_t* stray_pre; // [kJoinGrid + 3]
  void* scan_scratch;  // block totals of the two-launch scans (2 x 4096 x 8 bytes)
  void* desc;          // work item descriptors, 32 bytes each
  uint64_t desc_cap;
  size_t bytes;
};

int tuple_bytes(int key_bytes);
// Two passes over more partitions than the combined histogram holds.
bool deep_plan(const PartParams& pp);
uint32_t choose_nseg(int64_t n);
size_t layout_join(int64_t nR, int64_t nS, int key_bytes, const PartParams& pp, int base, int64_t slice,
                   char* ws, Layout* L);

// Launchers (return hipError_t of the launch).
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
hipError_t launch_fill_u64(uint64_t* p, uint64_t v, int64_t n, hipStream_t s);
hipError_t launch_hist(const Layout& L, int key_bytes, hipStream_t s);
hipError_t launch_colscan(const Layout& L, hipStream_t s);
hipError_t launch_offsets(const Layout& L, hipStream_t s);
hipError_t launch_pass1(const Layout& L, int key_bytes, hipStream_t s);
hipError_t launch_pass2(const Layout& L, int key_bytes, hipStream_t s);
// Deep plans: pass-2 digit counts per item over pass 1's output (h2, ipre),
// their scan over each pass-1 partition's items and the fine totals (tot).
hipError_t launch_hist2(const Layout& L, int key_bytes, hipStream_t s);
hipError_t launch_segscan(const Layout& L, hipStream_t s);
// Pass-2 items of a deep plan: at most this many per table.
uint64_t deep_item_capacity(int64_t n, int key_bytes, const PartParams& pp);
hipError_t launch_join_plan(const Layout& L, int base, int64_t slice, hipStream_t s);
// digit_sentinel: the tuples were partitioned here by their low key bits
// (false for caller-partitioned tuples, which may not be).
hipError_t launch_join(const Layout& L, int key_bytes, int base, int64_t slice, int unique, pj_pair* out,
                       int64_t capacity, hipStream_t s, bool digit_sentinel = true);
hipError_t launch_defrag(const Layout& L, pj_pair* out, int64_t capacity, int64_t* d_count, hipStream_t s);

// Generic join over externally partitioned tuples (phase API).
hipError_t launch_join_generic(const void* rt, const uint32_t* roff, const void* st, const uint32_t* soff,
                               int key_bytes, int B, int base, int64_t slice, int unique, pj_pair* out,
                               int64_t capacity, int64_t* d_count, char* ws, hipStream_t s);
size_t join_generic_ws_bytes(int B);

// Build-table slots that kJoinLdsBytes leaves next to the stage of one probe
// round; a build chunk is at most half of it.  The table itself has the
// largest power-of-two bucket count within it (cap_buckets).
int join_table_capacity(int key_bytes);

// log2 of the bucket count of the table of a build chunk of c tuples: the
// power of two giving at least 4c slots (occupancy <= 1/4), at most
// cap_buckets.  bs = slots per bucket.  A chunk is at most table_cap / 2 - 1
// tuples and cap_buckets * bs > table_cap / 2, so every probe walk ends at an
// empty slot.
__host__ __device__ inline uint32_t join_bucket_log2(uint32_t c, uint32_t bs, uint32_t cap_buckets) {
  const uint32_t want = (4u * c + bs - 1u) / bs;  // buckets for 4c slots
  uint32_t lg = want <= 1u ? 0u : 32u - uint32_t(__builtin_clz(want - 1u));
  const uint32_t cap_lg = 31u - uint32_t(__builtin_clz(cap_buckets));
  return lg < cap_lg ? lg : cap_lg;
}

}  // namespace pj
