// pj_partition.hip -- radix partitioning kernels for gfx950.
//
// Pipeline per table (build R and probe S are processed in the same launches,
// blockIdx.y = table):
//   k_hist     one read of the keys; per-segment fine histogram of
//              key & (P-1) in LDS (and its fold to pass-1 digits), written as a
//              row of a [segment][bin] matrix.  Replaces big_histogram +
//              retrieve_histogram_of_pass_one + histogram_pass_one/two
//              (reference src/partition_kernels.cuh:85-194,328-406).
//   k_colscan  exclusive scan of every matrix column over segments: the exact
//              write offset of each (segment, digit), so the scatter passes need
//              no global atomics (the reference claims ranges with atomicAdd,
//              src/partition_kernels.cuh:302-306,540-543).
//   k_offsets  exclusive scan of the bin totals: final partition starts
//              (cub::DeviceScan::ExclusiveSum, src/partitioned_join.cuh:86,152).
//   k_pass1    one workgroup per segment: tiles of the segment are ranked by
//              digit with LDS atomics, reordered through an LDS staging tile and
//              written as whole 128-byte lines (unfinished lines are carried in
//              LDS to the next tile) to the workgroup's private, contiguous
//              range of every partition (partition_pass_one, :196-326).
//   k_pass2    pass-1 partition p1 is split at segment-group boundaries into
//              items whose per-digit output starts are known from the matrix;
//              the number of groups follows p1's size (skew), and the items are
//              cut into one equal run per XCD (partition_pass_two, :408-563
//              with handle_skewed_data's block assignment,
//              src/skewed_data_handler.cuh:58-91).
// Segment s of a table is processed on XCD s / (nseg/8) by both the histogram
// and pass 1 (workgroup b runs on XCD b % 8).
#include <cmath>
#include <type_traits>

#include "pj_device.h"
#include "pj_internal.h"

namespace pj {

// ---------------------------------------------------------------------------
// Fill kernels (graph-safe replacements for hipMemsetAsync).
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t v, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

static int fill_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return int(g);
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t* p, uint64_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Key loading: 4 consecutive keys starting at i (i is a multiple of 4); low 32
// bits only.  Lanes past `hi` get valid = false.
// ---------------------------------------------------------------------------
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kXscanBlock = 256;
constexpr int kXscanPer = 16;
constexpr int kXscanTile = kXscanBlock * kXscanPer;

template <int KB, bool VEC>
__device__ __forceinline__ void load4_low(const void* keys, uint32_t i, uint32_t hi, uint32_t low[4],
                                          bool valid[4]) {
  if (KB == 4) {
    const uint32_t* k = static_cast<const uint32_t*>(keys);
    if (VEC && i + 4 <= hi) {
      const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(k + i));
      low[0] = v.x; low[1] = v.y; low[2] = v.z; low[3] = v.w;
      valid[0] = valid[1] = valid[2] = valid[3] = true;
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        valid[c] = i + c < hi;
        low[c] = valid[c] ? __builtin_nontemporal_load(k + i + c) : 0u;
      }
    }
  } else {
    const uint64_t* k = static_cast<const uint64_t*>(keys);
    if (VEC && i + 4 <= hi) {
      const u32x4 a = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(k + i));
      const u32x4 b = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(k + i + 2));
      low[0] = a.x; low[1] = a.z; low[2] = b.x; low[3] = b.z;
      valid[0] = valid[1] = valid[2] = valid[3] = true;
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        valid[c] = i + c < hi;
        low[c] = valid[c] ? uint32_t(__builtin_nontemporal_load(k + i + c)) : 0u;
      }
    }
  }
}

// Sum over the 64 lanes of a wave (all lanes active), in every lane.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += uint32_t(__shfl_xor(int(v), o, kWave));
  return v;
}

// ---------------------------------------------------------------------------
// k_hist: per-segment fine histogram.
//   PACKED == false: P <= 32768 u32 LDS bins.
//   PACKED == true : P == 65536, two 16-bit bins per LDS word; the segment is
//   counted in epochs of at most 61440 keys (no field can overflow), after each
//   epoch the fields are folded into per-thread register accumulators.
// ---------------------------------------------------------------------------
template <int KB, bool PACKED, bool VEC>
__global__ __launch_bounds__(kHistBlock) void k_hist(PartTable t0, PartTable t1, PartParams pp) {
  constexpr int BLOCK = kHistBlock;
  // Rounds of 4 keys per thread loaded together before any is counted, so
  // each wave keeps BATCH loads in flight, and the next batch is loaded while
  // the current one is counted (two register sets; one round at a time
  // leaves the loop latency-bound at a fraction of HBM bandwidth).
  constexpr uint32_t BATCH = 4;
  constexpr uint32_t kEpochBatches = 3;  // packed bins: 3 * 4 * 4096 = 49152 keys < 65536 per epoch
  const PartTable T = blockIdx.y ? t1 : t0;
  const uint32_t b = blockIdx.x;
  if (b >= T.nseg) return;
  const uint32_t S = T.nseg / kNumXcd;
  const uint32_t seg = (b % kNumXcd) * S + b / kNumXcd;
  const uint32_t lo = seg_bound(T.n, T.nseg, seg), hi = seg_bound(T.n, T.nseg, seg + 1);
  const int P = pp.P;
  const uint32_t pmask = uint32_t(P - 1);
  const bool two_pass = pp.passes == 2;
  const int F1 = pp.F1, b2 = pp.b2;

  extern __shared__ uint32_t s_mem[];
  uint32_t* s_bins = s_mem;
  uint32_t* s_coarse = s_mem + (PACKED ? P / 2 : P);
  const int tid = threadIdx.x;
  const uint32_t nrounds = (hi - lo + 4 * BLOCK - 1) / (4 * BLOCK);
  const uint32_t nfull = VEC ? (hi - lo) / (4 * BLOCK * BATCH) : 0;  // batches without a bounds check

  auto count = [&](uint32_t low) {
    const uint32_t d = (low >> pp.shift0) & pmask;
    if (PACKED) atomicAdd(&s_bins[d >> 1], 1u << ((d & 1u) << 4));
    else atomicAdd(&s_bins[d], 1u);
  };
  // The wave's hot bin (same-address LDS atomics serialise): when
  // kHeavyLanes or more lanes of a batch's first key row fall in it, each key
  // row of the batch counts its lanes in that bin with one add; otherwise the
  // batch counts key by key and lane 0's first bin becomes the candidate for
  // the next batch.  Under a heavy key the candidate settles on its bin
  // within a few batches and keeps it, whichever lanes hold the key.
  uint32_t hot = 0xFFFFFFFFu;
  auto count_hot = [&](uint32_t low, uint32_t h) {
    const uint32_t d = (low >> pp.shift0) & pmask;
    const uint64_t same = __ballot(d == h);
    const uint32_t inc = d != h ? 1u : (mbcnt(same) == 0u ? uint32_t(__popcll(same)) : 0u);
    if (inc) {
      if (PACKED) atomicAdd(&s_bins[d >> 1], inc << ((d & 1u) << 4));
      else atomicAdd(&s_bins[d], inc);
    }
  };
  // Batch f: rounds f * BATCH .. f * BATCH + BATCH - 1, all in range.
  auto load_batch = [&](uint32_t f, uint32_t (&low)[BATCH][4]) {
#pragma unroll
    for (uint32_t u = 0; u < BATCH; ++u) {
      const uint32_t i = lo + ((f * BATCH + u) * BLOCK + tid) * 4;
      if (KB == 4) {
        const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(static_cast<const uint32_t*>(T.keys) + i));
        low[u][0] = v.x; low[u][1] = v.y; low[u][2] = v.z; low[u][3] = v.w;
      } else {
        const uint64_t* k = static_cast<const uint64_t*>(T.keys) + i;
        const u32x4 a = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(k));
        const u32x4 c = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(k + 2));
        low[u][0] = a.x; low[u][1] = a.z; low[u][2] = c.x; low[u][3] = c.z;
      }
    }
  };
  auto count_low = [&](const uint32_t (&low)[BATCH][4]) {
    const uint32_t d = (low[0][0] >> pp.shift0) & pmask;
    if (__popcll(__ballot(d == hot)) >= kHeavyLanes) {
#pragma unroll
      for (uint32_t u = 0; u < BATCH; ++u)
#pragma unroll
        for (int c = 0; c < 4; ++c) count_hot(low[u][c], hot);
    } else {
      hot = uint32_t(__builtin_amdgcn_readfirstlane(int(d)));
#pragma unroll
      for (uint32_t u = 0; u < BATCH; ++u)
#pragma unroll
        for (int c = 0; c < 4; ++c) count(low[u][c]);
    }
  };
  auto count_batch = [&](uint32_t f) {
    uint32_t low[BATCH][4];
    load_batch(f, low);
    count_low(low);
  };
  auto count_round = [&](uint32_t r) {
    uint32_t low[4];
    bool valid[4];
    load4_low<KB, VEC>(T.keys, lo + (r * BLOCK + tid) * 4, hi, low, valid);
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (valid[c]) count(low[c]);
  };

  if (!PACKED) {
    for (int i = tid; i < P; i += BLOCK) s_bins[i] = 0;
    if (two_pass)
      for (int i = tid; i < F1; i += BLOCK) s_coarse[i] = 0;
    lds_barrier();
    for (uint32_t f = 0; f < nfull; ++f) count_batch(f);
    for (uint32_t r = nfull * BATCH; r < nrounds; ++r) count_round(r);
    lds_barrier();
    uint32_t* row = T.mat + size_t(seg) * P;
    for (int i = tid; i < P; i += BLOCK) {
      const uint32_t v = s_bins[i];
      row[i] = v;
      if (two_pass) {
        if (b2 >= 6) {
          // a wave's 64 consecutive bins share one pass-1 bin: one add for
          // them (64 same-address LDS atomics serialise)
          const uint32_t w = wave_sum(v);
          if (lane_id() == 0 && w) atomicAdd(&s_coarse[i >> b2], w);
        } else if (v) {
          atomicAdd(&s_coarse[i >> b2], v);
        }
      }
    }
  } else {
    constexpr int WPT = 32;  // LDS words per thread: 32768 words / 1024 threads
    uint32_t acc[2 * WPT];
#pragma unroll
    for (int k = 0; k < 2 * WPT; ++k) acc[k] = 0;
#pragma unroll
    for (int k = 0; k < WPT; ++k) s_bins[tid + BLOCK * k] = 0;
    if (two_pass)
      for (int i = tid; i < F1; i += BLOCK) s_coarse[i] = 0;
    lds_barrier();
    // Folds the 16-bit fields into the per-thread accumulators and clears them.
    auto fold = [&]() {
      lds_barrier();
#pragma unroll
      for (int k = 0; k < WPT; ++k) {
        const uint32_t w = s_bins[tid + BLOCK * k];
        acc[2 * k] += w & 0xffffu;
        acc[2 * k + 1] += w >> 16;
        s_bins[tid + BLOCK * k] = 0;
      }
      lds_barrier();
    };
    // Full batches, pipelined: batch f + 1 loads while batch f is counted
    // (the last iteration reloads the last batch rather than branch, so the
    // loads have fixed counts and the waits stay exact); a fold closes every
    // kEpochBatches batches and the last one.
    if (nfull > 0) {
      uint32_t la[BATCH][4], lb[BATCH][4];
      load_batch(0, la);
      uint32_t f = 0;
      for (; f + 1 < nfull; f += 2) {
        load_batch(f + 1, lb);
        count_low(la);
        if ((f + 1) % kEpochBatches == 0) fold();
        load_batch(min(f + 2, nfull - 1), la);
        count_low(lb);
        if ((f + 2) % kEpochBatches == 0) fold();
      }
      if (f < nfull) count_low(la);
      fold();
    }
    constexpr uint32_t kEpochRounds = kEpochBatches * BATCH;
    for (uint32_t r0 = nfull * BATCH; r0 < nrounds; r0 += kEpochRounds) {
      const uint32_t r1 = min(nrounds, r0 + kEpochRounds);
      for (uint32_t r = r0; r < r1; ++r) count_round(r);
      fold();
    }
    uint2* row = reinterpret_cast<uint2*>(T.mat + size_t(seg) * P);
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
      row[tid + BLOCK * k] = make_uint2(acc[2 * k], acc[2 * k + 1]);
      if (two_pass) {
        const uint32_t bin = 2u * (tid + BLOCK * k);
        if (b2 >= 7) {  // a wave's 128 consecutive bins share one pass-1 bin
          const uint32_t w = wave_sum(acc[2 * k] + acc[2 * k + 1]);
          if (lane_id() == 0 && w) atomicAdd(&s_coarse[bin >> b2], w);
        } else {
          if (acc[2 * k]) atomicAdd(&s_coarse[bin >> b2], acc[2 * k]);
          if (acc[2 * k + 1]) atomicAdd(&s_coarse[(bin + 1) >> b2], acc[2 * k + 1]);
        }
      }
    }
  }
  if (two_pass) {
    lds_barrier();
    uint32_t* crow = T.matc + size_t(seg) * F1;
    for (int i = tid; i < F1; i += BLOCK) crow[i] = s_coarse[i];
  }
}

hipError_t launch_hist(const Layout& L, int key_bytes, hipStream_t s) {
  const PartParams& pp = L.pp;
  const bool packed = pp.P > 32768;
  const uint32_t nseg = L.t[0].nseg > L.t[1].nseg ? L.t[0].nseg : L.t[1].nseg;
  const bool vec = (reinterpret_cast<uintptr_t>(L.t[0].keys) % 16 == 0) &&
                   (L.ntables < 2 || reinterpret_cast<uintptr_t>(L.t[1].keys) % 16 == 0);
  const size_t lds = size_t(packed ? pp.P / 2 : pp.P) * 4 + (pp.passes == 2 ? size_t(pp.F1) * 4 : 0);
  dim3 grid(nseg, L.ntables), block(kHistBlock);
#define PJ_HIST(KB, PK, V) hipLaunchKernelGGL((k_hist<KB, PK, V>), grid, block, lds, s, L.t[0], L.t[1], pp)
  if (key_bytes == 4) {
    if (packed) { if (vec) PJ_HIST(4, true, true); else PJ_HIST(4, true, false); }
    else { if (vec) PJ_HIST(4, false, true); else PJ_HIST(4, false, false); }
  } else {
    if (packed) { if (vec) PJ_HIST(8, true, true); else PJ_HIST(8, true, false); }
    else { if (vec) PJ_HIST(8, false, true); else PJ_HIST(8, false, false); }
  }
#undef PJ_HIST
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// k_colscan: exclusive scan of each column over the rows, in place; column
// totals to tot.  One thread per column; jobs along blockIdx.y.
// ---------------------------------------------------------------------------
struct ColJob {
  uint32_t* mat;
  uint32_t* tot;
  uint32_t cols, rows;
};
struct ColJobs {
  ColJob j[4];
};

__global__ __launch_bounds__(256) void k_colscan(ColJobs jobs) {
  const ColJob J = jobs.j[blockIdx.y];
  const uint32_t q = blockIdx.x * 256u + threadIdx.x;
  if (J.mat == nullptr || q >= J.cols) return;
  uint32_t acc = 0;
  uint32_t r = 0;
  for (; r + 8 <= J.rows; r += 8) {
    uint32_t v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = J.mat[size_t(r + u) * J.cols + q];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      J.mat[size_t(r + u) * J.cols + q] = acc;
      acc += v[u];
    }
  }
  for (; r < J.rows; ++r) {
    const uint32_t v = J.mat[size_t(r) * J.cols + q];
    J.mat[size_t(r) * J.cols + q] = acc;
    acc += v;
  }
  J.tot[q] = acc;
}

hipError_t launch_colscan(const Layout& L, hipStream_t s) {
  ColJobs jobs = {};
  uint32_t maxcols = 0;
  int njobs = 0;
  for (int t = 0; t < L.ntables; ++t) jobs.j[njobs++] = ColJob{L.t[t].mat, L.t[t].tot, uint32_t(L.pp.P), L.t[t].nseg};
  if (L.pp.passes == 2)
    for (int t = 0; t < L.ntables; ++t)
      jobs.j[njobs++] = ColJob{L.t[t].matc, L.t[t].totc, uint32_t(L.pp.F1), L.t[t].nseg};
  maxcols = uint32_t(L.pp.P);
  hipLaunchKernelGGL(k_colscan, dim3((maxcols + 255) / 256, njobs), dim3(256), 0, s, jobs);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Two-launch exclusive scan of per-index values (partition offsets, join work
// list).  k_xscan_local: each block scans kXscanTile consecutive values
// (kXscanPer contiguous values per thread) and stores its total;
// k_xscan_apply adds the sum of the preceding block totals and lets the value
// type record the grand total.  blockIdx.y selects the table.
// ---------------------------------------------------------------------------
template <typename T, class V>
__global__ __launch_bounds__(kXscanBlock) void k_xscan_local(V v, int64_t n, T* bsum) {
  __shared__ T s_tmp[kXscanBlock / kWave + 1];
  const int y = blockIdx.y;
  const int64_t base = int64_t(blockIdx.x) * kXscanTile + int64_t(threadIdx.x) * kXscanPer;
  T vals[kXscanPer];
  T sum = 0;
#pragma unroll
  for (int k = 0; k < kXscanPer; ++k) {
    vals[k] = base + k < n ? T(v.value(y, base + k)) : T(0);
    sum += vals[k];
  }
  T total;
  T run;
  if constexpr (sizeof(T) == 8) run = block_exclusive_scan64<kXscanBlock>(sum, s_tmp, &total);
  else run = block_exclusive_scan<kXscanBlock>(sum, s_tmp, &total);
  T* out = v.out[y];
#pragma unroll
  for (int k = 0; k < kXscanPer; ++k) {
    if (base + k < n) out[base + k] = run;
    run += vals[k];
  }
  if (threadIdx.x == 0) bsum[size_t(y) * gridDim.x + blockIdx.x] = total;
}

template <typename T, class V>
__global__ __launch_bounds__(kXscanBlock) void k_xscan_apply(V v, int64_t n, const T* bsum) {
  __shared__ T s_red[kXscanBlock / kWave];
  const int y = blockIdx.y;
  const int nb = gridDim.x;
  T pre = 0;  // sum of the preceding block totals (nb is small)
  for (int b = threadIdx.x; b < int(blockIdx.x); b += kXscanBlock) pre += bsum[size_t(y) * nb + b];
#pragma unroll
  for (int dlt = kWave / 2; dlt > 0; dlt >>= 1) pre += __shfl_xor(pre, dlt, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) s_red[threadIdx.x / kWave] = pre;
  __syncthreads();
  T offset = 0;
#pragma unroll
  for (int w = 0; w < kXscanBlock / kWave; ++w) offset += s_red[w];
  T* out = v.out[y];
  const int64_t base = int64_t(blockIdx.x) * kXscanTile;
  for (int k = threadIdx.x; k < kXscanTile; k += kXscanBlock)
    if (base + k < n) out[base + k] += offset;
  if (blockIdx.x == nb - 1 && threadIdx.x == 0) v.finish(y, n, offset + bsum[size_t(y) * nb + blockIdx.x]);
}

template <typename T, class V>
hipError_t launch_xscan(const V& v, int64_t n, int ntab, T* bsum, hipStream_t s) {
  const int nb = int((n + kXscanTile - 1) / kXscanTile);
  hipLaunchKernelGGL((k_xscan_local<T, V>), dim3(nb, ntab), dim3(kXscanBlock), 0, s, v, n, bsum);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_xscan_apply<T, V>), dim3(nb, ntab), dim3(kXscanBlock), 0, s, v, n, bsum);
  return hipGetLastError();
}

// off[q] = exclusive scan of tot over q; off[P] = n.
struct OffsetValues {
  const uint32_t* tot[2];
  uint32_t* out[2];
  uint32_t rows[2];
  __device__ uint32_t value(int y, int64_t i) const { return tot[y][i]; }
  __device__ void finish(int y, int64_t n, uint32_t) const { out[y][n] = rows[y]; }
};

hipError_t launch_offsets(const Layout& L, hipStream_t s) {
  OffsetValues v{};
  for (int t = 0; t < L.ntables; ++t) {
    v.tot[t] = L.t[t].tot;
    v.out[t] = L.t[t].off;
    v.rows[t] = L.t[t].n;
  }
  return launch_xscan<uint32_t>(v, L.pp.P, L.ntables, reinterpret_cast<uint32_t*>(L.scan_scratch), s);
}

// item_scan[q] = exclusive scan over partitions of ceil(|R_q|/base) * ceil(|S_q|/slice)
// (0 when either side is empty); item_scan[P] = total; resets the output chunk counter.
struct ItemValues {
  const uint32_t* roff;
  const uint32_t* soff;
  uint64_t* out[2];
  uint32_t* chunk_ctr;
  FastDiv base, slice;  // magic-number division by the build chunk and the probe slice
  __device__ uint64_t value(int, int64_t q) const {
    const uint32_t nr = roff[q + 1] - roff[q], ns = soff[q + 1] - soff[q];  // < 2^31
    if (!nr || !ns) return 0ull;
    return uint64_t(fastdiv_div(int32_t(nr - 1), base) + 1) * uint64_t(fastdiv_div(int32_t(ns - 1), slice) + 1);
  }
  __device__ void finish(int, int64_t n, uint64_t total) const {
    out[0][n] = total;
    *chunk_ctr = 0;
  }
};

hipError_t launch_join_plan(const Layout& L, int base, int64_t slice, hipStream_t s) {
  ItemValues v{};
  v.roff = L.t[0].off;
  v.soff = L.t[1].off;
  v.out[0] = L.item_scan;
  v.out[1] = L.item_scan;
  v.chunk_ctr = L.chunk_ctr;
  v.base = fastdiv_make(base);
  v.slice = fastdiv_make(int32_t(slice));
  return launch_xscan<uint64_t>(v, L.pp.P, 1, reinterpret_cast<uint64_t*>(L.scan_scratch), s);
}

// ---------------------------------------------------------------------------
// Shared tile scatter: E tuples per thread already in registers (digit
// 0xFFFFFFFF marks an empty item) are ranked with LDS atomics (rank_tile), the
// tile histogram is scanned and the tile staged in LDS in digit order
// (scan_stage_tile), and the staged tile is written out in coalesced runs
// (flush_full / flush_tile).  The next tile's loads are issued between the
// ranking and the scan, when the fewest registers are live.  Tile t is
// flushed at the start of tile t + 1, before the loads of tile t + 2: the
// memory counter completes in order, so the wait for those loads finds the
// stores long complete instead of waiting for them at once.  Two count arrays
// alternate between tiles: the one for tile t + 1 is cleared while tile t is
// staged, so no barrier of its own is needed; the first barrier of tile t + 1
// (after its ranking) orders the flush's LDS reads before the scan rewrites
// s_delta and the stage.
// ---------------------------------------------------------------------------
// Ranks of a tile's items within their digits, from LDS atomics on the
// cleared s_cnt: pk = digit << 16 | rank (digits < 4096, ranks < 65536), or
// 0xFFFFFFFF for an empty item.
template <int E>
__device__ __forceinline__ void rank_tile(const uint32_t (&dg)[E], uint32_t* s_cnt, uint32_t (&pk)[E]) {
  // Under a heavy key many lanes of a row share one digit, and same-address
  // LDS atomics serialise.  When kHeavyLanes or more lanes of the wave's
  // first row share lane 0's digit, each row's lanes with lane 0's digit take
  // their ranks from one atomic add (Zipf(1.5) probes, 16M x 256M: 6.29 ->
  // 4.60 ms).  Otherwise -- uniform digits -- the atomics issue back to back.
  const uint32_t f0 = uint32_t(__builtin_amdgcn_readfirstlane(int(dg[0])));
  if (f0 == 0xFFFFFFFFu || __popcll(__ballot(dg[0] == f0)) < kHeavyLanes) {
#pragma unroll
    for (int e = 0; e < E; ++e)
      pk[e] = dg[e] != 0xFFFFFFFFu ? (dg[e] << 16) | atomicAdd(&s_cnt[dg[e]], 1u) : 0xFFFFFFFFu;
    return;
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const uint32_t d0 = uint32_t(__builtin_amdgcn_readfirstlane(int(dg[e])));
    const uint64_t same = __ballot(dg[e] == d0);
    if (d0 != 0xFFFFFFFFu && __popcll(same) >= kHeavyLanes) {
      uint32_t base = 0;
      if (dg[e] == d0 && mbcnt(same) == 0u) base = atomicAdd(&s_cnt[d0], uint32_t(__popcll(same)));
      base = uint32_t(__builtin_amdgcn_readfirstlane(int(base)));
      if (dg[e] == d0) pk[e] = (d0 << 16) | (base + mbcnt(same));
      else pk[e] = dg[e] != 0xFFFFFFFFu ? (dg[e] << 16) | atomicAdd(&s_cnt[dg[e]], 1u) : 0xFFFFFFFFu;
    } else {
      pk[e] = dg[e] != 0xFFFFFFFFu ? (dg[e] << 16) | atomicAdd(&s_cnt[dg[e]], 1u) : 0xFFFFFFFFu;
    }
  }
}

// The scatter kernels hold and move tuples as vectors of the tuple's bytes
// (uint2: key, payload; four words: key low, key high, payload, 0): a struct
// value merged over branches or copied between register sets is kept in
// scratch memory, whose accesses join the in-order memory counter.  16-byte
// tuples are a native 4-word vector rather than HIP's uint4, a union whose
// copies go through 12-byte memory temporaries in scratch.  The key's low
// 32 bits are word x in both.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <int KB>
struct ScatterTup {
  using type = std::conditional_t<KB == 4, uint2, u32x4>;
  static_assert(sizeof(type) == sizeof(typename KeyT<KB>::tup_t), "tuple and vector sizes differ");
  __device__ __forceinline__ static type make(typename KeyT<KB>::key_t k, int32_t p) {
    if constexpr (KB == 4) return make_uint2(uint32_t(k), uint32_t(p));
    else return u32x4{uint32_t(k), uint32_t(uint64_t(k) >> 32), uint32_t(p), 0u};
  }
};

// Write combining.  HBM3E writes whole 128-byte lines: a store that covers
// part of a line makes the memory controller read the rest first, and a
// scatter whose runs start and end mid-line moves at ~3.7 TB/s where whole
// lines move at ~5 TB/s on MI355X.  So each digit keeps the unfinished last
// line of its run in LDS (line[d * LT + position % LT]) and a tile's flush
// stores only lines that are complete; a carried line goes out when a later
// tile completes it, or with the range's last tile.  A lane with nothing to
// store writes its sink slot instead (a private, never-read line of the
// workgroup), so every thread issues a fixed number of stores and the
// compiler's waits for the next tile's loads stay exact.  CF is the largest
// digit count of the instantiation (0: no carry).
template <int KB>
struct Carry {
  using vec_t = typename ScatterTup<KB>::type;
  static constexpr int LT = 128 / int(sizeof(vec_t));  // tuples per 128-byte line
  vec_t* line;            // [CF][LT] line images
  const uint32_t* start;  // [F] first position of the range per digit
  vec_t* out;             // the pass's output table
  vec_t* sink;            // this lane's sink slot
  uint32_t n_out;         // its rows (a carried position past them is never stored)
};

// Per workgroup slot and lane, 16 bytes each; only ever written.
constexpr int kSinkSlots = 1024;
static __device__ uint4 g_scatter_sink[kSinkSlots * kWave];

template <int KB>
__device__ __forceinline__ typename ScatterTup<KB>::type* sink_slot() {
  const uint32_t wg = (blockIdx.y * gridDim.x + blockIdx.x) % uint32_t(kSinkSlots);
  return reinterpret_cast<typename ScatterTup<KB>::type*>(g_scatter_sink + size_t(wg) * kWave + lane_id());
}

// The carried lines that a tile completes are read into registers at the end
// of the tile's staging and stored by the tile's flush, together with the
// rest of their lines: written apart, a half line could leave the L2 before
// its other half arrives and reach HBM as a partial line after all.
template <int KB, int BLOCK, int CF>
struct CarryRegs {
  static constexpr int NR = CF > 0 ? (CF * Carry<KB>::LT + BLOCK - 1) / BLOCK : 1;
  typename Carry<KB>::vec_t v[NR];
  uint32_t pos[NR];  // destination row, or 0xFFFFFFFF: nothing to store
};

// After a tile's scan (s_cur advanced by the tile's counts s_cnt): the carried
// line of each digit whose run the tile completes (or every one, for the last
// tile) -- its positions from the range start or the line start, whichever
// is later, up to the old cursor.
template <int KB, int BLOCK, int CF>
__device__ __forceinline__ void carry_read(int F, const uint32_t* s_cnt, const uint32_t* s_cur, const Carry<KB>& cy,
                                           bool last, CarryRegs<KB, BLOCK, CF>& cr) {
  constexpr uint32_t LT = uint32_t(Carry<KB>::LT), NS = uint32_t(CF) * LT;
  // The thread index through an opaque move: the slot addresses below are
  // then recomputed per tile (a few adds) instead of being hoisted out of
  // the tile loop, where they would be spilled, and a spill reload waits for
  // every outstanding load and store.
  uint32_t tid;
  __asm__ volatile("v_mov_b32 %0, %1" : "=v"(tid) : "v"(uint32_t(threadIdx.x)));
#pragma unroll
  for (int r = 0; r < CarryRegs<KB, BLOCK, CF>::NR; ++r) {
    const uint32_t slot = min(uint32_t(r) * BLOCK + tid, NS - 1), d = slot / LT, k = slot % LT;
    const uint32_t dd = d < uint32_t(F) ? d : uint32_t(F) - 1;
    const uint32_t cur = s_cur[dd], old = cur - s_cnt[dd], l0 = old & ~(LT - 1), pos = l0 + k;
    const bool done = last || (cur & ~(LT - 1)) > l0;
    const bool valid = uint32_t(r) * BLOCK + tid < NS && d < uint32_t(F) && done && pos < old && pos >= cy.start[dd] &&
                       pos < cy.n_out;
    cr.v[r] = cy.line[slot];
    cr.pos[r] = valid ? pos : 0xFFFFFFFFu;
  }
}

// NR stores per thread, unconditional (the sink takes the empty ones).
template <int KB, int BLOCK, int CF>
__device__ __forceinline__ void carry_write(const Carry<KB>& cy, const CarryRegs<KB, BLOCK, CF>& cr) {
#pragma unroll
  for (int r = 0; r < CarryRegs<KB, BLOCK, CF>::NR; ++r)
    *(cr.pos[r] != 0xFFFFFFFFu ? cy.out + cr.pos[r] : cy.sink) = cr.v[r];
}

// Inclusive prefix sum over the 64 lanes of a wave with DPP moves: row shifts
// by 1, 2, 4 and 8 scan each row of 16 lanes, then row broadcasts of lanes 15
// and 31 carry the row totals into the later rows (lanes with no source add 0).
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
  x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x111, 0xf, 0xf, false));  // row_shr:1
  x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x112, 0xf, 0xf, false));  // row_shr:2
  x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x114, 0xf, 0xf, false));  // row_shr:4
  x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x118, 0xf, 0xf, false));  // row_shr:8
  x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x142, 0xa, 0xf, false));  // row_bcast:15 -> rows 1, 3
  x += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x143, 0xc, 0xf, false));  // row_bcast:31 -> rows 2, 3
  return x;
}

// After rank_tile: the first wave scans the F digit counts (F / 64 per lane,
// no barrier), advances the per-digit global cursors s_cur and records s_delta
// (global minus staged position per digit); then the tile is staged in digit
// order while s_cnt_next is cleared for the next tile.  Three barriers; ends
// with one.
template <int KB, int BLOCK, int E, int CF>
__device__ __forceinline__ void scan_stage_tile(const typename ScatterTup<KB>::type (&tup)[E], const uint32_t (&pk)[E],
                                                int F, typename ScatterTup<KB>::type* s_stage, const uint32_t* s_cnt,
                                                uint32_t* s_cnt_next, uint32_t* s_base, uint32_t* s_cur,
                                                uint32_t* s_delta, const Carry<KB>& cy, bool last,
                                                CarryRegs<KB, BLOCK, CF>& cr) {
  const int tid = threadIdx.x;
  lds_barrier();
  if (tid < kWave) {
    const int DPL = F > kWave ? F / kWave : 1;  // digits per lane
    const int d0 = tid * DPL;
    if constexpr (CF > 0) {
      // at most CF / 64 digits per lane: unrolled, so the LDS reads issue together
      constexpr int DMAX = CF > kWave ? CF / kWave : 1;
      uint32_t c[DMAX], cur[DMAX], local = 0;
#pragma unroll
      for (int k = 0; k < DMAX; ++k) {
        const bool ok = k < DPL && d0 + k < F;
        c[k] = ok ? s_cnt[d0 + k] : 0u;
        cur[k] = ok ? s_cur[d0 + k] : 0u;
        local += c[k];
      }
      uint32_t run = wave_inclusive_scan(local) - local;
#pragma unroll
      for (int k = 0; k < DMAX; ++k)
        if (k < DPL && d0 + k < F) {
          s_base[d0 + k] = run;
          s_delta[d0 + k] = cur[k] - run;
          s_cur[d0 + k] = cur[k] + c[k];
          run += c[k];
        }
    } else {
      uint32_t local = 0;
      for (int k = 0; k < DPL; ++k)
        if (d0 + k < F) local += s_cnt[d0 + k];
      uint32_t run = wave_inclusive_scan(local) - local;
      for (int k = 0; k < DPL; ++k) {
        const int d = d0 + k;
        if (d < F) {
          const uint32_t c = s_cnt[d];
          s_base[d] = run;
          s_delta[d] = s_cur[d] - run;
          s_cur[d] += c;
          run += c;
        }
      }
    }
  }
  lds_barrier();
  for (int d = tid; d < F; d += BLOCK) s_cnt_next[d] = 0;
#pragma unroll
  for (int e = 0; e < E; ++e)
    if (pk[e] != 0xFFFFFFFFu) s_stage[s_base[pk[e] >> 16] + (pk[e] & 0xFFFFu)] = tup[e];
  if constexpr (CF > 0) carry_read<KB, BLOCK, CF>(F, s_cnt, s_cur, cy, last, cr);
  lds_barrier();
}

// Flush of a full tile: every thread stores exactly E tuples, unconditionally,
// so the compiler knows how many memory operations follow the next tile's
// loads and waits for those alone.  A destination past the table (impossible
// with exact offsets) is clamped to its last row rather than faulting.
// With a carry (CF > 0), tuples in the unfinished last line of their digit's
// run go to the line image instead (and the lane stores to its sink); after
// the range's last tile (last) nothing is held back.
template <int KB, int BLOCK, int E, int CF>
__device__ __forceinline__ void flush_full(int shift, uint32_t mask, const typename ScatterTup<KB>::type* s_stage,
                                           const uint32_t* s_delta, const uint32_t* s_cur,
                                           typename ScatterTup<KB>::type* out, uint32_t n_out, const Carry<KB>& cy,
                                           bool last, const CarryRegs<KB, BLOCK, CF>& cr) {
  using V = typename ScatterTup<KB>::type;
  constexpr uint32_t LT = uint32_t(Carry<KB>::LT);
  const int tid = threadIdx.x;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const uint32_t j = uint32_t(e) * BLOCK + tid;
    const V t = s_stage[j];
    const uint32_t dg = digit_of(t.x, shift, mask);
    const uint32_t p = s_delta[dg] + j;
    V* dst = out + (p < n_out ? p : n_out - 1);
    if constexpr (CF > 0) {
      const uint32_t cut = last ? 0xFFFFFFFFu : (s_cur[dg] & ~(LT - 1));
      const bool keep = p >= cut;
      if (keep) cy.line[dg * LT + (p & (LT - 1))] = t;
      dst = keep ? cy.sink : dst;
    }
    *dst = t;
  }
  if constexpr (CF > 0) carry_write<KB, BLOCK, CF>(cy, cr);
}

// Flush of a partial tile of len tuples (the last of a range, so nothing is
// carried), with the carried lines the tile's staging read.
template <int KB, int BLOCK, int E, int CF>
__device__ __forceinline__ void flush_tile(uint32_t len, int shift, uint32_t mask,
                                           const typename ScatterTup<KB>::type* s_stage, const uint32_t* s_delta,
                                           typename ScatterTup<KB>::type* out, uint32_t n_out, const Carry<KB>& cy,
                                           const CarryRegs<KB, BLOCK, CF>& cr) {
  const int tid = threadIdx.x;
  for (uint32_t j = tid; j < len; j += BLOCK) {
    const typename ScatterTup<KB>::type t = s_stage[j];
    const uint32_t d = s_delta[digit_of(t.x, shift, mask)] + j;
    if (d < n_out) out[d] = t;  // always in range for exact offsets; guards the table
  }
  if constexpr (CF > 0) carry_write<KB, BLOCK, CF>(cy, cr);
}

// LDS layout of the scatter kernels: stage[TILE], then two count arrays,
// base, cur and delta, F words each; with a carry also start[F] and the line
// images [CF][128 bytes].
template <int KB, int TILE, int CF>
__host__ __device__ constexpr size_t scatter_lds_bytes(int F) {
  return size_t(TILE) * (KB == 4 ? 8 : 16) + size_t(F) * (CF > 0 ? 24 : 20) + size_t(CF) * 128;
}

// Tile and carry of a scatter pass over F digits, so that stage, counts and
// line images fit 160 KB of LDS beside pass 2's item prefix (E = TILE / 1024
// per thread, a multiple of 4 in pass 1, whose loads are 16 bytes wide):
//   4-byte keys  F <= 128: 16K + carry; F <= 256: 12K (pass 1) or 14K (pass 2)
//                + carry; 512: 8K + carry; 1024: 16K; more: 8K;
//   8-byte keys  (16-byte tuples) F <= 64: 8K + carry (pass 2), F <= 128: 8K +
//                carry (pass 1); F <= 256: 6K (pass 1) or 7K (pass 2) + carry;
//                512: 4K + carry; 1024: 8K; more: 4K.
// Pass 2 after 4096 pass-1 partitions: the item prefix takes 16 KB, so 128,
// 256 and 1024 digits drop to the next instantiated tile that fits (no carry
// for 128 and 256 digits: 16K or 8K tuples; 1024 digits: 8K or 4K), and more
// than 2048 digits leave room for 7K (4-byte keys) or 3K tuples.
struct ScatterCfg {
  int tile, cf;
};
inline ScatterCfg scatter_cfg(int key_bytes, int F, bool pass2, int F1 = 0) {
  if (pass2 && F1 > 2048) {
    if (F > 2048) return {key_bytes == 4 ? 7168 : 3072, 0};
    if (F > 64 && F <= 256) return {key_bytes == 4 ? 16384 : 8192, 0};
    if (F == 1024) return {key_bytes == 4 ? 8192 : 4096, 0};
  }
  if (key_bytes == 4) {
    if (F <= 64) return {16384, 64};
    if (F <= 128) return {16384, 128};
    if (F <= 256) return {pass2 ? 14336 : 12288, 256};
    if (F == 512) return {8192, 512};
    if (F == 1024) return {16384, 0};
    return {8192, 0};
  }
  if (F <= (pass2 ? 64 : 128)) return {8192, pass2 ? 64 : 128};
  if (F <= 256) return {pass2 ? 7168 : 6144, 256};
  if (F == 512) return {4096, 512};
  if (F == 1024) return {8192, 0};
  return {4096, 0};
}

// ---------------------------------------------------------------------------
// k_pass1: one workgroup per segment; digit (key >> b2) & (F1 - 1).  Full
// tiles load without any condition, so the next tile's loads stay in flight
// until their registers are used; only the segment's partial last tile loads
// under bounds checks.  PAY0 / PAY1: table 0 / 1 has a payload column (else
// the payload is the row id).
// ---------------------------------------------------------------------------
template <int KB, bool VEC, bool PAY, int BLOCK, int TILE, int CF>
__device__ __forceinline__ void pass1_body(const PartTable& T, const PartParams& pp, uint32_t b) {
  using KTr = KeyT<KB>;
  using key_t = typename KTr::key_t;
  using tup_t = typename ScatterTup<KB>::type;  // the tuple as a vector
  constexpr int E = TILE / BLOCK;
  constexpr int G = KB == 4 ? 4 : 2;  // keys per 16-byte load
  static_assert(E % G == 0, "tile must give each thread whole 16-byte groups of keys");
  const uint32_t S = T.nseg / kNumXcd;
  const uint32_t seg = (b % kNumXcd) * S + b / kNumXcd;
  const uint32_t lo = seg_bound(T.n, T.nseg, seg), hi = seg_bound(T.n, T.nseg, seg + 1);
  const int F = pp.F1, shift = pp.shift0 + pp.b2;
  const uint32_t mask = uint32_t(F - 1);

  extern __shared__ __align__(16) uint8_t s_raw[];
  tup_t* s_stage = reinterpret_cast<tup_t*>(s_raw);
  uint32_t* s_cnt = reinterpret_cast<uint32_t*>(s_raw + size_t(TILE) * sizeof(tup_t));  // this tile's counts
  uint32_t* s_cnt_next = s_cnt + F;                                                     // the next tile's
  uint32_t* s_base = s_cnt_next + F;
  uint32_t* s_cur = s_base + F;
  uint32_t* s_delta = s_cur + F;
  uint32_t* s_start = s_delta + F;
  tup_t* out = static_cast<tup_t*>(T.buf1);
  const Carry<KB> cy{reinterpret_cast<tup_t*>(s_start + F), s_start, out, sink_slot<KB>(), T.n};
  const int tid = threadIdx.x;

  const uint32_t* rel = (pp.passes == 2) ? T.matc + size_t(seg) * F : T.mat + size_t(seg) * F;
  for (int d = tid; d < F; d += BLOCK) {
    s_cur[d] = T.off[size_t(d) << pp.b2] + rel[d];
    if (CF > 0) s_start[d] = s_cur[d];
    s_cnt[d] = 0;
  }
  lds_barrier();
  auto swap_counts = [&]() {
    uint32_t* t = s_cnt;
    s_cnt = s_cnt_next;
    s_cnt_next = t;
  };
  const key_t* keys = static_cast<const key_t*>(T.keys);

  // Item g * G + c of a thread is row tile_lo + (g * BLOCK + tid) * G + c:
  // one 16-byte load of keys per group (payloads: G * 4 bytes).
  auto load_full = [&](uint32_t tile_lo, key_t (&key)[E], int32_t (&pay)[E]) {
#pragma unroll
    for (int g = 0; g < E / G; ++g) {
      const uint32_t i = tile_lo + (uint32_t(g) * BLOCK + tid) * G;
      if (VEC) {
        const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(keys + i));
        if constexpr (KB == 4) {
          key[g * 4 + 0] = key_t(v.x); key[g * 4 + 1] = key_t(v.y);
          key[g * 4 + 2] = key_t(v.z); key[g * 4 + 3] = key_t(v.w);
        } else {
          key[g * 2 + 0] = key_t(uint64_t(v.x) | (uint64_t(v.y) << 32));
          key[g * 2 + 1] = key_t(uint64_t(v.z) | (uint64_t(v.w) << 32));
        }
        if (PAY) {
          if constexpr (G == 4) {
            const u32x4 q = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(T.pay + i));
            pay[g * 4 + 0] = int32_t(q.x); pay[g * 4 + 1] = int32_t(q.y);
            pay[g * 4 + 2] = int32_t(q.z); pay[g * 4 + 3] = int32_t(q.w);
          } else {
            typedef uint32_t u32x2v __attribute__((ext_vector_type(2)));
            const u32x2v q = __builtin_nontemporal_load(reinterpret_cast<const u32x2v*>(T.pay + i));
            pay[g * 2 + 0] = int32_t(q.x); pay[g * 2 + 1] = int32_t(q.y);
          }
        }
      } else {
#pragma unroll
        for (int c = 0; c < G; ++c) {
          key[g * G + c] = __builtin_nontemporal_load(keys + i + c);
          if (PAY) pay[g * G + c] = __builtin_nontemporal_load(T.pay + i + c);
        }
      }
    }
  };  // without a payload column pay is left alone: the row id is formed when staging
  auto load_tail = [&](uint32_t tile_lo, key_t (&key)[E], int32_t (&pay)[E]) {
#pragma unroll
    for (int g = 0; g < E / G; ++g) {
      const uint32_t i = tile_lo + (uint32_t(g) * BLOCK + tid) * G;
#pragma unroll
      for (int c = 0; c < G; ++c) {
        const bool valid = i + c < hi;
        key[g * G + c] = valid ? keys[i + c] : key_t(0);
        pay[g * G + c] = PAY ? (valid ? T.pay[i + c] : 0) : int32_t(i + c);
      }
    }
  };
  auto digits = [&](uint32_t tile_lo, const key_t (&key)[E], uint32_t (&dg)[E]) {
#pragma unroll
    for (int g = 0; g < E / G; ++g) {
      const uint32_t i = tile_lo + (uint32_t(g) * BLOCK + tid) * G;
#pragma unroll
      for (int c = 0; c < G; ++c)
        dg[g * G + c] = i + c < hi ? digit_of(KTr::low32(key[g * G + c]), shift, mask) : 0xFFFFFFFFu;
    }
  };

  // Full tiles.  Tile 0 is peeled so that every later iteration flushes the
  // previous tile (E stores per thread) and loads the next one (clamped to the
  // last full tile, which the final iteration reloads) without conditions:
  // with fixed counts the compiler's waits for a tile's loads leave the
  // younger stores and loads in flight.
  const uint32_t nfull = (hi - lo) / uint32_t(TILE);
  const bool no_tail = lo + nfull * uint32_t(TILE) == hi;
  CarryRegs<KB, BLOCK, CF> cr;  // carried lines read by a tile's staging, stored by its flush
  // One full tile held in (ck, cp): flush the previous tile (flush_prev),
  // rank this one, issue the next tile's loads into (nk, np), scan and stage.
  // The caller alternates two register sets, so a loaded tile is never copied
  // (a copy would wait for the loads).
  auto full_tile = [&](uint32_t t, auto flush_prev, const key_t (&ck)[E], const int32_t (&cp)[E], key_t (&nk)[E],
                       int32_t (&np)[E]) {
    if (decltype(flush_prev)::value)
      flush_full<KB, BLOCK, E, CF>(shift, mask, s_stage, s_delta, s_cur, out, T.n, cy, false, cr);
    uint32_t pk[E];
    {
      uint32_t dg[E];
#pragma unroll
      for (int e = 0; e < E; ++e) dg[e] = digit_of(KTr::low32(ck[e]), shift, mask);
      rank_tile<E>(dg, s_cnt, pk);
    }
    const uint32_t tn = t + 1 < nfull ? t + 1 : t;
    load_full(lo + tn * uint32_t(TILE), nk, np);  // next tile in flight during the scan and staging
    tup_t tup[E];
    const uint32_t tile_lo = lo + t * uint32_t(TILE);
#pragma unroll
    for (int e = 0; e < E; ++e)
      tup[e] = ScatterTup<KB>::make(ck[e], PAY ? cp[e] : int32_t(tile_lo + (uint32_t(e / G) * BLOCK + tid) * G + (e % G)));
    scan_stage_tile<KB, BLOCK, E, CF>(tup, pk, F, s_stage, s_cnt, s_cnt_next, s_base, s_cur, s_delta, cy,
                                      no_tail && t + 1 == nfull, cr);
    swap_counts();
  };
  key_t key[E];
  int32_t pay[E];
  if (nfull > 0) {
    key_t key2[E];
    int32_t pay2[E];
    load_full(lo, key, pay);
    full_tile(0u, std::false_type{}, key, pay, key2, pay2);
    uint32_t t = 1;
    for (; t + 1 < nfull; t += 2) {
      full_tile(t, std::true_type{}, key2, pay2, key, pay);
      full_tile(t + 1, std::true_type{}, key, pay, key2, pay2);
    }
    if (t < nfull) full_tile(t, std::true_type{}, key2, pay2, key, pay);
    flush_full<KB, BLOCK, E, CF>(shift, mask, s_stage, s_delta, s_cur, out, T.n, cy, no_tail, cr);
  }
  // The partial last tile.
  const uint32_t tail_lo = lo + nfull * uint32_t(TILE);
  if (tail_lo < hi) {
    load_tail(tail_lo, key, pay);
    uint32_t dg[E], pk[E];
    digits(tail_lo, key, dg);
    rank_tile<E>(dg, s_cnt, pk);
    tup_t tup[E];
#pragma unroll
    for (int e = 0; e < E; ++e) tup[e] = ScatterTup<KB>::make(key[e], pay[e]);
    scan_stage_tile<KB, BLOCK, E, CF>(tup, pk, F, s_stage, s_cnt, s_cnt_next, s_base, s_cur, s_delta, cy, true, cr);
    flush_tile<KB, BLOCK, E, CF>(hi - tail_lo, shift, mask, s_stage, s_delta, out, T.n, cy, cr);
  }
}

template <int KB, bool VEC, bool PAY0, bool PAY1, int BLOCK, int TILE, int CF>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_pass1(PartTable t0, PartTable t1, PartParams pp) {
  const uint32_t b = blockIdx.x;
  if (blockIdx.y == 0) {
    if (b < t0.nseg) pass1_body<KB, VEC, PAY0, BLOCK, TILE, CF>(t0, pp, b);
  } else {
    if (b < t1.nseg) pass1_body<KB, VEC, PAY1, BLOCK, TILE, CF>(t1, pp, b);
  }
}

template <int KB, bool VEC, bool PAY0, bool PAY1, int TILE, int CF>
static hipError_t pass1_launch_t(const Layout& L, hipStream_t s) {
  const uint32_t nseg = L.t[0].nseg > L.t[1].nseg ? L.t[0].nseg : L.t[1].nseg;
  const size_t lds = scatter_lds_bytes<KB, TILE, CF>(L.pp.F1);
  hipLaunchKernelGGL((k_pass1<KB, VEC, PAY0, PAY1, kScatterBlock, TILE, CF>), dim3(nseg, L.ntables),
                     dim3(kScatterBlock), lds, s, L.t[0], L.t[1], L.pp);
  return hipGetLastError();
}

// Unaligned key columns (rare) take the plain tiles without a carry.
template <int KB, bool VEC, bool PAY0, bool PAY1>
static hipError_t pass1_launch(const Layout& L, hipStream_t s) {
  constexpr int BIG = KB == 4 ? kScatterTile4 : kScatterTile8;
  if (!VEC) {
    if (L.pp.F1 <= 1024) return pass1_launch_t<KB, VEC, PAY0, PAY1, BIG, 0>(L, s);
    return pass1_launch_t<KB, VEC, PAY0, PAY1, BIG / 2, 0>(L, s);
  }
  const ScatterCfg c = scatter_cfg(KB, L.pp.F1, false);
#define PJ_P1(T_, C_) \
  if (c.tile == T_ && c.cf == C_) return pass1_launch_t<KB, VEC, PAY0, PAY1, T_, C_>(L, s)
  if constexpr (KB == 4) {
    PJ_P1(16384, 64);
    PJ_P1(16384, 128);
    PJ_P1(12288, 256);
    PJ_P1(8192, 512);
    PJ_P1(16384, 0);
    PJ_P1(8192, 0);
  } else {
    PJ_P1(8192, 128);
    PJ_P1(6144, 256);
    PJ_P1(4096, 512);
    PJ_P1(8192, 0);
    PJ_P1(4096, 0);
  }
#undef PJ_P1
  return hipErrorInvalidValue;
}

template <int KB, bool VEC>
static hipError_t pass1_dispatch(const Layout& L, hipStream_t s) {
  const bool p0 = L.t[0].pay != nullptr;
  const bool p1 = L.ntables > 1 && L.t[1].pay != nullptr;
  if (p0) return p1 ? pass1_launch<KB, VEC, true, true>(L, s) : pass1_launch<KB, VEC, true, false>(L, s);
  return p1 ? pass1_launch<KB, VEC, false, true>(L, s) : pass1_launch<KB, VEC, false, false>(L, s);
}

static bool aligned16(const void* p) { return p == nullptr || (reinterpret_cast<uintptr_t>(p) % 16) == 0; }

hipError_t launch_pass1(const Layout& L, int key_bytes, hipStream_t s) {
  bool vec = aligned16(L.t[0].keys) && aligned16(L.t[0].pay);
  if (L.ntables > 1) vec = vec && aligned16(L.t[1].keys) && aligned16(L.t[1].pay);
  if (key_bytes == 4) return vec ? pass1_dispatch<4, true>(L, s) : pass1_dispatch<4, false>(L, s);
  return vec ? pass1_dispatch<8, true>(L, s) : pass1_dispatch<8, false>(L, s);
}

// ---------------------------------------------------------------------------
// Pass-2 work items.  Pass-1 partition p1 is split into ng(p1) items, numbered
// in XCD-major partition order r = (p1 % 8) * PX + p1 / 8, PX = ceil(F1 / 8):
//   combined-histogram plans: ng = clamp(|p1| / chunk, 1, nseg) groups of
//     consecutive segments (0 for an empty p1), whose per-digit output starts
//     the [segment][digit] matrix gives;
//   deep plans: ng = ceil(|p1| / chunk) consecutive ranges of `chunk` tuples,
//     whose per-digit counts k_hist2 takes over pass 1's output.
// s_pre[r] counts the items before rank r, s_pre[RN] all of them.  Ends with
// a barrier.
//
// With s_run (pass 2), the items are also cut into 8 runs of about equal
// tuple counts, one per XCD: an item's start is taken as t(r) + g * n / ng
// (t(r) the tuples of the ranks before r, n and ng its partition's tuples
// and items), and s_run[0] / s_run[1] receive the first item whose start
// reaches x * N / 8 / (x + 1) * N / 8 (N all tuples; x = xcd).  Equal item
// counts per run would give the XCD that holds a heavy key's partition all
// of its items, each up to 256 times the size of an ordinary one.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t p2_rank_to_p1(uint32_t r, uint32_t PX) { return (r % PX) * kNumXcd + r / PX; }

template <int BLOCK, bool DEEP>
__device__ __forceinline__ void pass2_item_prefix(const PartTable& T, int F1, int b2, uint32_t chunk, uint32_t* s_pre,
                                                  uint32_t* s_tmp, uint32_t* s_tmp2 = nullptr, uint32_t xcd = 0,
                                                  uint32_t* s_run = nullptr) {
  const uint32_t PX = uint32_t(F1 + kNumXcd - 1) / kNumXcd, RN = PX * kNumXcd;
  constexpr int RPT = (kMaxRadix + BLOCK - 1) / BLOCK;  // ranks per thread
  const int tid = threadIdx.x;
  uint32_t ng[RPT], nt[RPT], sum = 0, tsum = 0;
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const uint32_t r = uint32_t(tid) * RPT + i, p1 = r < RN ? p2_rank_to_p1(r, PX) : 0xFFFFFFFFu;
    ng[i] = 0;
    nt[i] = 0;
    if (p1 < uint32_t(F1)) {
      if constexpr (DEEP) {
        nt[i] = T.offc[p1 + 1] - T.offc[p1];
        ng[i] = (nt[i] + chunk - 1) / chunk;
      } else {
        nt[i] = T.off[(p1 + 1) << b2] - T.off[p1 << b2];
        ng[i] = nt[i] == 0 ? 0u : min(T.nseg, max(1u, nt[i] / chunk));
      }
    }
    sum += ng[i];
    tsum += nt[i];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan<BLOCK>(sum, s_tmp, &total);
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const uint32_t r = uint32_t(tid) * RPT + i;
    if (r < RN) s_pre[r] = run;
    run += ng[i];
  }
  if (tid == 0) s_pre[RN] = total;
  if (s_run != nullptr) {
    // the run bounds: the rank r < RN with t(r) < A <= t(r) + n (the last
    // rank if none) holds the first item whose start reaches A
    uint32_t ttotal;
    uint32_t trun = block_exclusive_scan<BLOCK>(tsum, s_tmp2, &ttotal);
    uint32_t irun = run - sum;
    const uint32_t A[2] = {uint32_t(uint64_t(xcd) * ttotal / kNumXcd),
                           uint32_t(uint64_t(xcd + 1) * ttotal / kNumXcd)};
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const uint32_t r = uint32_t(tid) * RPT + i;
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        if (r < RN && trun < A[b] && (r + 1 == RN || trun + nt[i] >= A[b])) {
          const uint64_t c = nt[i] == 0 ? 0ull : (uint64_t(A[b] - trun) * ng[i] + nt[i] - 1) / nt[i];
          s_run[b] = irun + uint32_t(c < ng[i] ? c : ng[i]);
        }
      }
      trun += nt[i];
      irun += ng[i];
    }
    // no rank has t(r) < 0: a bound at A = 0 is item 0 (A = N > 0 finds the
    // last non-empty rank, and with it all the items)
    if (tid == 0) {
      if (A[0] == 0) s_run[0] = 0;
      if (A[1] == 0) s_run[1] = 0;
    }
  }
  lds_barrier();
}

// ---------------------------------------------------------------------------
// k_pass2: items (p1, segment group or range) -> split by digit key & (F2 - 1).
// ---------------------------------------------------------------------------
template <int KB, int BLOCK, int TILE, int CF, bool DEEP>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_pass2(PartTable t0, PartTable t1,
                                                                                        PartParams pp, uint32_t g0) {
  using tup_t = typename ScatterTup<KB>::type;  // the tuple as a vector
  constexpr int E = TILE / BLOCK;
  // Workgroups [0, g0) serve table 0 and [g0, grid) table 1, in proportion
  // to the tables' sizes; g0 is a multiple of 8, so a workgroup's XCD is its
  // index within its table modulo 8.
  const bool second = blockIdx.x >= g0;
  const PartTable T = second ? t1 : t0;
  const uint32_t local = second ? blockIdx.x - g0 : blockIdx.x;
  const uint32_t groups = second ? gridDim.x - g0 : g0;
  const int x = int(local % kNumXcd), j = int(local / kNumXcd), G = int(groups / kNumXcd);
  const int F1 = pp.F1, F2 = pp.F2, b2 = pp.b2, P = pp.P;
  const uint32_t mask = uint32_t(F2 - 1);
  const uint32_t nseg = T.nseg;

  extern __shared__ __align__(16) uint8_t s_raw[];
  tup_t* s_stage = reinterpret_cast<tup_t*>(s_raw);
  uint32_t* s_cnt = reinterpret_cast<uint32_t*>(s_raw + size_t(TILE) * sizeof(tup_t));  // this tile's counts
  uint32_t* s_cnt_next = s_cnt + F2;                                                    // the next tile's
  uint32_t* s_base = s_cnt_next + F2;
  uint32_t* s_cur = s_base + F2;
  uint32_t* s_delta = s_cur + F2;
  uint32_t* s_start = s_delta + F2;
  tup_t* s_line = reinterpret_cast<tup_t*>(s_start + (CF > 0 ? F2 : 0));
  // Work items (pass2_item_prefix): an item holds about kPass2ItemTiles full
  // tiles whatever the skew.
  const uint32_t PX = uint32_t(F1 + kNumXcd - 1) / kNumXcd, RN = PX * kNumXcd;
  uint32_t* s_pre = reinterpret_cast<uint32_t*>(s_line + CF * Carry<KB>::LT);  // RN + 1 entries
  uint32_t* s_tmp = s_pre + RN + 1;   // BLOCK / 64 + 1 entries
  uint32_t* s_tmp2 = s_tmp + BLOCK / kWave + 1;  // BLOCK / 64 + 1 entries
  uint32_t* s_run = s_tmp2 + BLOCK / kWave + 1;  // 2 entries
  const int tid = threadIdx.x;
  const uint32_t chunk = uint32_t(kPass2ItemTiles) * uint32_t(TILE);
  auto p1_of = [&](uint32_t r) { return p2_rank_to_p1(r, PX); };
  for (int d = tid; d < F2; d += BLOCK) s_cnt[d] = 0;
  pass2_item_prefix<BLOCK, DEEP>(T, F1, b2, chunk, s_pre, s_tmp, s_tmp2, uint32_t(x), s_run);
  // This workgroup's items (LDS values read as scalars: every lane holds the
  // same): the items, in rank order, are cut into 8 runs of about equal tuple
  // counts and the workgroups of XCD x stride over run x.  With equal XCD
  // shares run x is exactly the partitions p1 = x mod 8, so a partition's
  // partly written output lines meet in one L2; a heavy partition's excess
  // moves to the neighbouring XCDs (16M x 256M Zipf 1.0: pass 2 1.47 -> 1.19
  // ms, against 1.15 ms with no XCD placement at all, which costs 2.4 % on
  // uniform keys).
  auto pre = [&](uint32_t r) { return uint32_t(__builtin_amdgcn_readfirstlane(s_pre[r])); };
  const uint32_t k_begin = uint32_t(__builtin_amdgcn_readfirstlane(s_run[0])) + uint32_t(j);
  const uint32_t k_end = uint32_t(__builtin_amdgcn_readfirstlane(s_run[1]));
  auto swap_counts = [&]() {
    uint32_t* t = s_cnt;
    s_cnt = s_cnt_next;
    s_cnt_next = t;
  };
  const tup_t* in = static_cast<const tup_t*>(T.buf1);
  tup_t* out = static_cast<tup_t*>(T.buf2);
  const Carry<KB> cy{s_line, s_start, out, sink_slot<KB>(), T.n};

  auto digits = [&](const tup_t (&tv)[E], uint32_t tile_lo, uint32_t hi, uint32_t (&dg)[E]) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const bool valid = tile_lo + uint32_t(e) * BLOCK + tid < hi;
      dg[e] = valid ? digit_of(tv[e].x, pp.shift0, mask) : 0xFFFFFFFFu;
    }
  };
  for (uint32_t k = k_begin; k < k_end; k += uint32_t(G)) {
    uint32_t lo = 0, hi = RN;  // s_pre[lo] <= k < s_pre[hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (pre(mid) <= k) lo = mid; else hi = mid;
    }
    const int p1 = int(p1_of(lo));
    const uint32_t NG = pre(lo + 1) - pre(lo), g = k - pre(lo);
    const uint32_t q0 = uint32_t(p1) << b2;
    uint32_t in_lo, in_hi, seg_lo = 0;
    if constexpr (DEEP) {
      in_lo = T.offc[p1] + g * chunk;
      in_hi = min(in_lo + chunk, T.offc[p1 + 1]);
    } else {
      seg_lo = g * nseg / NG;
      const uint32_t seg_hi = (g + 1) * nseg / NG;
      const uint32_t base1 = T.off[q0];
      in_lo = base1 + T.matc[size_t(seg_lo) * F1 + p1];
      in_hi = seg_hi < nseg ? base1 + T.matc[size_t(seg_hi) * F1 + p1] : T.off[q0 + F2];
    }
    if (in_lo >= in_hi || in_hi > T.n) continue;
    // s_cur and s_start were last read by the previous tile's scan, before
    // its staging barrier (the flush after a range's last tile reads neither)
    for (int d = tid; d < F2; d += BLOCK) {
      const uint32_t c0 = DEEP ? T.off[q0 + d] + T.h2[size_t(k) * F2 + d]
                               : T.off[q0 + d] + T.mat[size_t(seg_lo) * P + q0 + d];
      s_cur[d] = c0;
      if (CF > 0) s_start[d] = c0;
    }
    // Full tiles, pipelined as in pass 1: tile 0 peeled, the next tile's
    // loads clamped to the last full tile, E stores per thread per flush.
    const uint32_t nfull = (in_hi - in_lo) / uint32_t(TILE);
    const bool no_tail = in_lo + nfull * uint32_t(TILE) == in_hi;
    CarryRegs<KB, BLOCK, CF> cr;  // carried lines read by a tile's staging, stored by its flush
    auto full_tile = [&](uint32_t t, auto flush_prev, const tup_t (&cur)[E], tup_t (&nxt)[E]) {
      if (decltype(flush_prev)::value)
        flush_full<KB, BLOCK, E, CF>(pp.shift0, mask, s_stage, s_delta, s_cur, out, T.n, cy, false, cr);
      uint32_t pk[E];
      {
        uint32_t dg[E];
#pragma unroll
        for (int e = 0; e < E; ++e) dg[e] = digit_of(cur[e].x, pp.shift0, mask);
        rank_tile<E>(dg, s_cnt, pk);
      }
      // Next tile in flight: the next full tile, or after the last one the
      // item's partial tail (indices clamped to the item, so the loads need
      // no condition and nothing is read twice).
      const uint32_t next_lo = in_lo + (t + 1) * uint32_t(TILE);
#pragma unroll
      for (int e = 0; e < E; ++e) nxt[e] = in[min(next_lo + uint32_t(e) * BLOCK + tid, in_hi - 1)];
      scan_stage_tile<KB, BLOCK, E, CF>(cur, pk, F2, s_stage, s_cnt, s_cnt_next, s_base, s_cur, s_delta, cy,
                                        no_tail && t + 1 == nfull, cr);
      swap_counts();
    };
    tup_t tv[E], tw[E];
    if (nfull > 0) {
#pragma unroll
      for (int e = 0; e < E; ++e) tv[e] = in[in_lo + uint32_t(e) * BLOCK + tid];
      full_tile(0u, std::false_type{}, tv, tw);
      uint32_t t = 1;
      for (; t + 1 < nfull; t += 2) {
        full_tile(t, std::true_type{}, tw, tv);
        full_tile(t + 1, std::true_type{}, tv, tw);
      }
      if (t < nfull) full_tile(t, std::true_type{}, tw, tv);
      flush_full<KB, BLOCK, E, CF>(pp.shift0, mask, s_stage, s_delta, s_cur, out, T.n, cy, no_tail, cr);
      // the prefetched tail is in tw after an odd number of full tiles
      if (nfull & 1u) {
#pragma unroll
        for (int e = 0; e < E; ++e) tv[e] = tw[e];
      }
    }
    // The partial last tile of the item.
    const uint32_t tail_lo = in_lo + nfull * uint32_t(TILE);
    if (tail_lo < in_hi) {
      if (nfull == 0) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const uint32_t i = tail_lo + uint32_t(e) * BLOCK + tid;
          if (i < in_hi) tv[e] = in[i];  // lanes past the end are marked empty by their digit
        }
      }
      uint32_t dg[E], pk[E];
      digits(tv, tail_lo, in_hi, dg);
      rank_tile<E>(dg, s_cnt, pk);
      scan_stage_tile<KB, BLOCK, E, CF>(tv, pk, F2, s_stage, s_cnt, s_cnt_next, s_base, s_cur, s_delta, cy, true, cr);
      swap_counts();
      flush_tile<KB, BLOCK, E, CF>(in_hi - tail_lo, pp.shift0, mask, s_stage, s_delta, out, T.n, cy, cr);
    }
  }
}

// Persistent pass-2 workgroups, kPass2WgPerXcd per XCD and table on average,
// split between the tables by their sizes (the work is proportional to them):
// the grid, and g0, the first workgroup of table 1 (a multiple of 8).
static void pass2_grid(const Layout& L, uint32_t* grid, uint32_t* g0) {
  const uint32_t per_xcd = uint32_t(kPass2WgPerXcd) * uint32_t(L.ntables);
  uint32_t x0 = per_xcd;
  if (L.ntables > 1) {
    const double n0 = double(L.t[0].n), n1 = double(L.t[1].n);
    const double share = n0 + n1 > 0 ? n0 / (n0 + n1) : 0.5;
    x0 = uint32_t(std::lround(share * per_xcd));
    if (x0 < 1) x0 = 1;
    if (x0 > per_xcd - 1) x0 = per_xcd - 1;
  }
  *g0 = x0 * uint32_t(kNumXcd);
  *grid = per_xcd * kNumXcd;
}

static uint32_t pass2_chunk(int key_bytes, const PartParams& pp) {
  return uint32_t(kPass2ItemTiles) * uint32_t(scatter_cfg(key_bytes, pp.F2, true, pp.F1).tile);
}

uint64_t deep_item_capacity(int64_t n, int key_bytes, const PartParams& pp) {
  return uint64_t(n) / pass2_chunk(key_bytes, pp) + uint64_t(pp.F1) + 1;
}

hipError_t launch_pass2(const Layout& L, int key_bytes, hipStream_t s) {
  uint32_t ngrid = 0, g0 = 0;
  pass2_grid(L, &ngrid, &g0);
  const bool deep = L.pp.B > kMaxCombinedBits;
  dim3 grid(ngrid), block(kScatterBlock);
  // the item prefix (one entry per pass-1 partition rounded up to 8, plus one)
  // and the scan's temporaries follow the scatter's LDS
  const size_t extra = (size_t((L.pp.F1 + kNumXcd - 1) / kNumXcd) * kNumXcd + 1 + 2 * (kScatterBlock / kWave + 1) + 2) * 4;
  const ScatterCfg c = scatter_cfg(key_bytes, L.pp.F2, true, L.pp.F1);
#define PJ_PASS2(KB_, TILE_, CF_)                                                                                   \
  if (c.tile == TILE_ && c.cf == CF_) {                                                                             \
    const size_t lds_ = scatter_lds_bytes<KB_, TILE_, CF_>(L.pp.F2) + extra;                                        \
    if (deep) hipLaunchKernelGGL((k_pass2<KB_, kScatterBlock, TILE_, CF_, true>), grid, block, lds_, s, L.t[0],    \
                                 L.t[1], L.pp, g0);                                                                 \
    else hipLaunchKernelGGL((k_pass2<KB_, kScatterBlock, TILE_, CF_, false>), grid, block, lds_, s, L.t[0],        \
                            L.t[1], L.pp, g0);                                                                      \
    return hipGetLastError();                                                                                       \
  }
  if (key_bytes == 4) {
    PJ_PASS2(4, 16384, 64)
    PJ_PASS2(4, 16384, 128)
    PJ_PASS2(4, 14336, 256)
    PJ_PASS2(4, 8192, 512)
    PJ_PASS2(4, 16384, 0)
    PJ_PASS2(4, 8192, 0)
    PJ_PASS2(4, 7168, 0)
  } else {
    PJ_PASS2(8, 8192, 64)
    PJ_PASS2(8, 7168, 256)
    PJ_PASS2(8, 4096, 512)
    PJ_PASS2(8, 8192, 0)
    PJ_PASS2(8, 4096, 0)
    PJ_PASS2(8, 3072, 0)
  }
#undef PJ_PASS2
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Deep plans (more partitions than the combined histogram holds): after
// pass 1, k_hist2 counts the pass-2 digits of every pass-2 item over pass 1's
// output (the reference's histogram_pass_two, src/partition_kernels.cuh:328-406,
// per item rather than per block), k_segscan turns the counts into each
// item's per-digit output starts within its pass-1 partition and the fine
// partition totals, and the offsets scan gives the final partition starts.
// ---------------------------------------------------------------------------
template <int KB, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_hist2(PartTable t0, PartTable t1, PartParams pp, uint32_t chunk,
                                                 uint32_t g0) {
  using tup_t = typename KeyT<KB>::tup_t;
  const bool second = blockIdx.x >= g0;
  const PartTable T = second ? t1 : t0;
  const uint32_t local = second ? blockIdx.x - g0 : blockIdx.x;
  const uint32_t groups = second ? gridDim.x - g0 : g0;
  const int F1 = pp.F1, F2 = pp.F2;
  const uint32_t mask = uint32_t(F2 - 1);
  const uint32_t PX = uint32_t(F1 + kNumXcd - 1) / kNumXcd, RN = PX * kNumXcd;
  extern __shared__ uint32_t s_mem[];
  uint32_t* s_bins = s_mem;            // F2
  uint32_t* s_pre = s_bins + F2;       // RN + 1
  uint32_t* s_tmp = s_pre + RN + 1;    // BLOCK / 64 + 1
  const int tid = threadIdx.x;
  pass2_item_prefix<BLOCK, true>(T, F1, pp.b2, chunk, s_pre, s_tmp);
  if (local == 0)
    for (uint32_t r = tid; r <= RN; r += BLOCK) T.ipre[r] = s_pre[r];
  const uint32_t n_items = s_pre[RN];
  const tup_t* in = static_cast<const tup_t*>(T.buf1);
  for (uint32_t k = local; k < n_items; k += groups) {
    uint32_t lo = 0, hi = RN;  // s_pre[lo] <= k < s_pre[hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_pre[mid] <= k) lo = mid; else hi = mid;
    }
    const uint32_t p1 = p2_rank_to_p1(lo, PX), g = k - s_pre[lo];
    const uint32_t in_lo = T.offc[p1] + g * chunk, in_hi = min(in_lo + chunk, T.offc[p1 + 1]);
    for (int d = tid; d < F2; d += BLOCK) s_bins[d] = 0;
    lds_barrier();
    for (uint32_t i = in_lo + tid; i < in_hi; i += BLOCK) atomicAdd(&s_bins[in[i].x & mask], 1u);
    lds_barrier();
    for (int d = tid; d < F2; d += BLOCK) T.h2[size_t(k) * F2 + d] = s_bins[d];
    lds_barrier();
  }
}

hipError_t launch_hist2(const Layout& L, int key_bytes, hipStream_t s) {
  uint32_t ngrid = 0, g0 = 0;
  pass2_grid(L, &ngrid, &g0);
  const uint32_t RN = uint32_t((L.pp.F1 + kNumXcd - 1) / kNumXcd) * kNumXcd;
  const size_t lds = (size_t(L.pp.F2) + RN + 1 + kScatterBlock / kWave + 1) * 4;
  const uint32_t chunk = pass2_chunk(key_bytes, L.pp);
  if (key_bytes == 4)
    hipLaunchKernelGGL((k_hist2<4, kScatterBlock>), dim3(ngrid), dim3(kScatterBlock), lds, s, L.t[0], L.t[1], L.pp,
                       chunk, g0);
  else
    hipLaunchKernelGGL((k_hist2<8, kScatterBlock>), dim3(ngrid), dim3(kScatterBlock), lds, s, L.t[0], L.t[1], L.pp,
                       chunk, g0);
  return hipGetLastError();
}

// One thread per fine partition q = p1 * F2 + d: the exclusive scan of digit
// d's counts over p1's items (in place) and their total, tot[q].
__global__ __launch_bounds__(256) void k_segscan(PartTable t0, PartTable t1, PartParams pp) {
  const PartTable T = blockIdx.y ? t1 : t0;
  const uint32_t q = blockIdx.x * 256u + threadIdx.x;
  if (q >= uint32_t(pp.P)) return;
  const uint32_t F2 = uint32_t(pp.F2), p1 = q >> pp.b2, d = q & (F2 - 1u);
  const uint32_t PX = uint32_t(pp.F1 + kNumXcd - 1) / kNumXcd;
  const uint32_t r = (p1 % kNumXcd) * PX + p1 / kNumXcd;
  uint32_t acc = 0;
  for (uint32_t k = T.ipre[r]; k < T.ipre[r + 1]; ++k) {
    const uint32_t v = T.h2[size_t(k) * F2 + d];
    T.h2[size_t(k) * F2 + d] = acc;
    acc += v;
  }
  T.tot[q] = acc;
}

hipError_t launch_segscan(const Layout& L, hipStream_t s) {
  hipLaunchKernelGGL(k_segscan, dim3((L.pp.P + 255) / 256, L.ntables), dim3(256), 0, s, L.t[0], L.t[1], L.pp);
  return hipGetLastError();
}

}  // namespace pj
