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
//              written as coalesced runs to the workgroup's private, contiguous
//              range of every partition (partition_pass_one, :196-326).
//   k_pass2    pass-1 partition p1 is split at segment-group boundaries into
//              items whose per-digit output starts are known from the matrix;
//              all items of one p1 run on one XCD so partial lines merge in its
//              L2 (partition_pass_two, :408-563 with handle_skewed_data's block
//              assignment, src/skewed_data_handler.cuh:58-91).
// Segment s of a table is processed on XCD s / (nseg/8) by both the histogram
// and pass 1 (workgroup b runs on XCD b % 8).
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

  if (!PACKED) {
    for (int i = tid; i < P; i += BLOCK) s_bins[i] = 0;
    if (two_pass)
      for (int i = tid; i < F1; i += BLOCK) s_coarse[i] = 0;
    lds_barrier();
    for (uint32_t r = 0; r < nrounds; ++r) {
      uint32_t low[4];
      bool valid[4];
      load4_low<KB, VEC>(T.keys, lo + (r * BLOCK + tid) * 4, hi, low, valid);
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (valid[c]) atomicAdd(&s_bins[(low[c] >> pp.shift0) & pmask], 1u);
    }
    lds_barrier();
    uint32_t* row = T.mat + size_t(seg) * P;
    for (int i = tid; i < P; i += BLOCK) {
      const uint32_t v = s_bins[i];
      row[i] = v;
      if (two_pass && v) atomicAdd(&s_coarse[i >> b2], v);
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
    constexpr uint32_t kEpochRounds = 15;  // 15 * 4096 = 61440 keys < 65536
    for (uint32_t r0 = 0; r0 < nrounds; r0 += kEpochRounds) {
      const uint32_t r1 = min(nrounds, r0 + kEpochRounds);
      for (uint32_t r = r0; r < r1; ++r) {
        uint32_t low[4];
        bool valid[4];
        load4_low<KB, VEC>(T.keys, lo + (r * BLOCK + tid) * 4, hi, low, valid);
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (valid[c]) {
            const uint32_t d = (low[c] >> pp.shift0) & pmask;
            atomicAdd(&s_bins[d >> 1], 1u << ((d & 1u) << 4));
          }
      }
      lds_barrier();
#pragma unroll
      for (int k = 0; k < WPT; ++k) {
        const uint32_t w = s_bins[tid + BLOCK * k];
        acc[2 * k] += w & 0xffffu;
        acc[2 * k + 1] += w >> 16;
        s_bins[tid + BLOCK * k] = 0;
      }
      lds_barrier();
    }
    uint2* row = reinterpret_cast<uint2*>(T.mat + size_t(seg) * P);
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
      row[tid + BLOCK * k] = make_uint2(acc[2 * k], acc[2 * k + 1]);
      if (two_pass) {
        const uint32_t bin = 2u * (tid + BLOCK * k);
        if (acc[2 * k]) atomicAdd(&s_coarse[bin >> b2], acc[2 * k]);
        if (acc[2 * k + 1]) atomicAdd(&s_coarse[(bin + 1) >> b2], acc[2 * k + 1]);
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
  uint32_t base;
  uint64_t slice;
  __device__ uint64_t value(int, int64_t q) const {
    const uint64_t nr = roff[q + 1] - roff[q], ns = soff[q + 1] - soff[q];
    return (nr && ns) ? ((nr + base - 1) / base) * ((ns + slice - 1) / slice) : 0ull;
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
  v.base = uint32_t(base);
  v.slice = uint64_t(slice);
  return launch_xscan<uint64_t>(v, L.pp.P, 1, reinterpret_cast<uint64_t*>(L.scan_scratch), s);
}

// ---------------------------------------------------------------------------
// Shared tile scatter: E items per thread already in registers (digit 0xFFFFFFFF
// marks an empty item).  Ranks by LDS atomics, scans the tile histogram,
// advances the per-digit global cursors s_cur, stages the tile in LDS in digit
// order and writes it out in coalesced runs.  Ends with an LDS barrier.
// ---------------------------------------------------------------------------
template <int KB, int BLOCK, int E>
__device__ __forceinline__ void scatter_tile(const typename KeyT<KB>::key_t (&key)[E], const int32_t (&pay)[E],
                                             const uint32_t (&dg)[E], uint32_t len, int F, int shift,
                                             uint32_t mask, typename KeyT<KB>::tup_t* s_stage, uint32_t* s_cnt,
                                             uint32_t* s_base, uint32_t* s_cur, uint32_t* s_delta, uint32_t* s_tmp,
                                             typename KeyT<KB>::tup_t* out, uint32_t n_out) {
  using KTr = KeyT<KB>;
  const int tid = threadIdx.x;
  uint32_t rk[E];
#pragma unroll
  for (int e = 0; e < E; ++e) rk[e] = dg[e] != 0xFFFFFFFFu ? atomicAdd(&s_cnt[dg[e]], 1u) : 0u;
  lds_barrier();
  // Exclusive scan of the F digit counts: each thread owns DPT consecutive digits.
  const int DPT = F > BLOCK ? F / BLOCK : 1;
  const int d0 = tid * DPT;
  uint32_t local = 0;
  for (int k = 0; k < DPT; ++k)
    if (d0 + k < F) local += s_cnt[d0 + k];
  uint32_t total;
  uint32_t run = block_exclusive_scan<BLOCK>(local, s_tmp, &total);
  for (int k = 0; k < DPT; ++k) {
    const int d = d0 + k;
    if (d < F) {
      const uint32_t c = s_cnt[d];
      s_base[d] = run;
      s_delta[d] = s_cur[d] - run;
      s_cur[d] += c;
      run += c;
    }
  }
  lds_barrier();
#pragma unroll
  for (int e = 0; e < E; ++e)
    if (dg[e] != 0xFFFFFFFFu) s_stage[s_base[dg[e]] + rk[e]] = KTr::make(key[e], pay[e]);
  lds_barrier();
  for (uint32_t j = tid; j < len; j += BLOCK) {
    const typename KTr::tup_t t = s_stage[j];
    const uint32_t d = digit_of(KTr::low32(KTr::key(t)), shift, mask);
    const uint32_t pos = s_delta[d] + j;
    if (pos < n_out) out[pos] = t;  // always true for exact offsets; guards the table bounds
  }
  lds_barrier();
}

// LDS layout of the scatter kernels: stage[TILE] then cnt, base, cur, delta [F] and tmp.
template <int KB, int TILE>
__host__ __device__ constexpr size_t scatter_lds_bytes(int F) {
  return size_t(TILE) * (KB == 4 ? 8 : 16) + size_t(F) * 16 + 64 * 4;
}

// ---------------------------------------------------------------------------
// k_pass1: one workgroup per segment; digit (key >> b2) & (F1 - 1).
// ---------------------------------------------------------------------------
template <int KB, bool VEC, int BLOCK, int TILE>
__global__ __launch_bounds__(BLOCK) void k_pass1(PartTable t0, PartTable t1, PartParams pp) {
  using KTr = KeyT<KB>;
  using key_t = typename KTr::key_t;
  using tup_t = typename KTr::tup_t;
  constexpr int E = TILE / BLOCK;
  static_assert(E % 4 == 0, "tile must give each thread a multiple of 4 items");
  const PartTable T = blockIdx.y ? t1 : t0;
  const uint32_t b = blockIdx.x;
  if (b >= T.nseg) return;
  const uint32_t S = T.nseg / kNumXcd;
  const uint32_t seg = (b % kNumXcd) * S + b / kNumXcd;
  const uint32_t lo = seg_bound(T.n, T.nseg, seg), hi = seg_bound(T.n, T.nseg, seg + 1);
  const int F = pp.F1, shift = pp.shift0 + pp.b2;
  const uint32_t mask = uint32_t(F - 1);
  const bool haspay = T.pay != nullptr;

  extern __shared__ __align__(16) uint8_t s_raw[];
  tup_t* s_stage = reinterpret_cast<tup_t*>(s_raw);
  uint32_t* s_cnt = reinterpret_cast<uint32_t*>(s_raw + size_t(TILE) * sizeof(tup_t));
  uint32_t* s_base = s_cnt + F;
  uint32_t* s_cur = s_base + F;
  uint32_t* s_delta = s_cur + F;
  uint32_t* s_tmp = s_delta + F;
  const int tid = threadIdx.x;

  const uint32_t* rel = (pp.passes == 2) ? T.matc + size_t(seg) * F : T.mat + size_t(seg) * F;
  for (int d = tid; d < F; d += BLOCK) s_cur[d] = T.off[size_t(d) << pp.b2] + rel[d];
  tup_t* out = static_cast<tup_t*>(T.buf1);
  const key_t* keys = static_cast<const key_t*>(T.keys);

  // Loads the 4-item groups of one tile into registers (keys and payloads).
  auto load_tile = [&](uint32_t tile_lo, key_t (&key)[E], int32_t (&pay)[E]) {
#pragma unroll
    for (int e4 = 0; e4 < E / 4; ++e4) {
      const uint32_t i = tile_lo + (uint32_t(e4) * BLOCK + tid) * 4;
      if (VEC && i + 4 <= hi) {
        if (KB == 4) {
          const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(keys + i));
          key[e4 * 4 + 0] = key_t(v.x); key[e4 * 4 + 1] = key_t(v.y);
          key[e4 * 4 + 2] = key_t(v.z); key[e4 * 4 + 3] = key_t(v.w);
        } else {
          const u32x4 a = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(keys + i));
          const u32x4 c = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(keys + i + 2));
          key[e4 * 4 + 0] = key_t(uint64_t(a.x) | (uint64_t(a.y) << 32));
          key[e4 * 4 + 1] = key_t(uint64_t(a.z) | (uint64_t(a.w) << 32));
          key[e4 * 4 + 2] = key_t(uint64_t(c.x) | (uint64_t(c.y) << 32));
          key[e4 * 4 + 3] = key_t(uint64_t(c.z) | (uint64_t(c.w) << 32));
        }
        if (haspay) {
          const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(T.pay + i));
          pay[e4 * 4 + 0] = int32_t(v.x); pay[e4 * 4 + 1] = int32_t(v.y);
          pay[e4 * 4 + 2] = int32_t(v.z); pay[e4 * 4 + 3] = int32_t(v.w);
        } else {
#pragma unroll
          for (int c = 0; c < 4; ++c) pay[e4 * 4 + c] = int32_t(i + c);
        }
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const bool valid = i + c < hi;
          key[e4 * 4 + c] = valid ? __builtin_nontemporal_load(keys + i + c) : key_t(0);
          pay[e4 * 4 + c] = haspay ? (valid ? __builtin_nontemporal_load(T.pay + i + c) : 0) : int32_t(i + c);
        }
      }
    }
  };
  key_t key[E];
  int32_t pay[E];
  if (lo < hi) load_tile(lo, key, pay);
  for (uint32_t tile_lo = lo; tile_lo < hi; tile_lo += TILE) {
    const uint32_t len = min(uint32_t(TILE), hi - tile_lo);
    // next tile in flight while this one is ranked, staged and written
    key_t nkey[E];
    int32_t npay[E];
    const uint32_t next_lo = tile_lo + TILE;
    if (next_lo < hi) load_tile(next_lo, nkey, npay);
    for (int d = tid; d < F; d += BLOCK) s_cnt[d] = 0;
    lds_barrier();
    uint32_t dg[E];
#pragma unroll
    for (int e4 = 0; e4 < E / 4; ++e4) {
      const uint32_t i = tile_lo + (uint32_t(e4) * BLOCK + tid) * 4;
#pragma unroll
      for (int c = 0; c < 4; ++c)
        dg[e4 * 4 + c] = i + c < hi ? digit_of(KTr::low32(key[e4 * 4 + c]), shift, mask) : 0xFFFFFFFFu;
    }
    scatter_tile<KB, BLOCK, E>(key, pay, dg, len, F, shift, mask, s_stage, s_cnt, s_base, s_cur, s_delta, s_tmp,
                               out, T.n);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      key[e] = nkey[e];
      pay[e] = npay[e];
    }
  }
}

template <int KB, bool VEC>
static hipError_t pass1_dispatch(const Layout& L, hipStream_t s) {
  constexpr int TILE = KB == 4 ? kScatterTile4 : kScatterTile8;
  const uint32_t nseg = L.t[0].nseg > L.t[1].nseg ? L.t[0].nseg : L.t[1].nseg;
  const size_t lds = scatter_lds_bytes<KB, TILE>(L.pp.F1);
  hipLaunchKernelGGL((k_pass1<KB, VEC, kScatterBlock, TILE>), dim3(nseg, L.ntables), dim3(kScatterBlock), lds, s,
                     L.t[0], L.t[1], L.pp);
  return hipGetLastError();
}

static bool aligned16(const void* p) { return p == nullptr || (reinterpret_cast<uintptr_t>(p) % 16) == 0; }

hipError_t launch_pass1(const Layout& L, int key_bytes, hipStream_t s) {
  bool vec = aligned16(L.t[0].keys) && aligned16(L.t[0].pay);
  if (L.ntables > 1) vec = vec && aligned16(L.t[1].keys) && aligned16(L.t[1].pay);
  if (key_bytes == 4) return vec ? pass1_dispatch<4, true>(L, s) : pass1_dispatch<4, false>(L, s);
  return vec ? pass1_dispatch<8, true>(L, s) : pass1_dispatch<8, false>(L, s);
}

// ---------------------------------------------------------------------------
// k_pass2: items (p1, segment group) -> split by digit key & (F2 - 1).
// ---------------------------------------------------------------------------
template <int KB, int BLOCK, int TILE>
__global__ __launch_bounds__(BLOCK) void k_pass2(PartTable t0, PartTable t1, PartParams pp) {
  using KTr = KeyT<KB>;
  using key_t = typename KTr::key_t;
  using tup_t = typename KTr::tup_t;
  constexpr int E = TILE / BLOCK;
  const PartTable T = blockIdx.y ? t1 : t0;
  const int x = blockIdx.x % kNumXcd, j = blockIdx.x / kNumXcd, G = gridDim.x / kNumXcd;
  const int F1 = pp.F1, F2 = pp.F2, b2 = pp.b2, P = pp.P;
  const uint32_t mask = uint32_t(F2 - 1);
  const uint32_t nseg = T.nseg;
  const uint32_t NG = nseg < uint32_t(kPass2Groups) ? nseg : uint32_t(kPass2Groups);
  const int n_p1 = F1 > x ? (F1 - x + kNumXcd - 1) / kNumXcd : 0;
  const uint32_t items = uint32_t(n_p1) * NG;

  extern __shared__ __align__(16) uint8_t s_raw[];
  tup_t* s_stage = reinterpret_cast<tup_t*>(s_raw);
  uint32_t* s_cnt = reinterpret_cast<uint32_t*>(s_raw + size_t(TILE) * sizeof(tup_t));
  uint32_t* s_base = s_cnt + F2;
  uint32_t* s_cur = s_base + F2;
  uint32_t* s_delta = s_cur + F2;
  uint32_t* s_tmp = s_delta + F2;
  const int tid = threadIdx.x;
  const tup_t* in = static_cast<const tup_t*>(T.buf1);
  tup_t* out = static_cast<tup_t*>(T.buf2);

  for (uint32_t k = j; k < items; k += G) {
    const int p1 = x + kNumXcd * int(k / NG);
    const uint32_t g = k % NG;
    const uint32_t seg_lo = g * nseg / NG, seg_hi = (g + 1) * nseg / NG;
    const uint32_t q0 = uint32_t(p1) << b2;
    const uint32_t base1 = T.off[q0];
    const uint32_t in_lo = base1 + T.matc[size_t(seg_lo) * F1 + p1];
    const uint32_t in_hi = seg_hi < nseg ? base1 + T.matc[size_t(seg_hi) * F1 + p1] : T.off[q0 + F2];
    if (in_lo >= in_hi || in_hi > T.n) continue;
    for (int d = tid; d < F2; d += BLOCK) s_cur[d] = T.off[q0 + d] + T.mat[size_t(seg_lo) * P + q0 + d];
    tup_t tv[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint32_t i = in_lo + uint32_t(e) * BLOCK + tid;
      if (i < in_hi) tv[e] = in[i];
    }
    for (uint32_t tile_lo = in_lo; tile_lo < in_hi; tile_lo += TILE) {
      const uint32_t len = min(uint32_t(TILE), in_hi - tile_lo);
      tup_t tn[E];
      const uint32_t next_lo = tile_lo + TILE;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const uint32_t i = next_lo + uint32_t(e) * BLOCK + tid;
        if (next_lo < in_hi && i < in_hi) tn[e] = in[i];
      }
      for (int d = tid; d < F2; d += BLOCK) s_cnt[d] = 0;
      lds_barrier();
      key_t key[E];
      int32_t pay[E];
      uint32_t dg[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const uint32_t i = tile_lo + uint32_t(e) * BLOCK + tid;
        const bool valid = i < in_hi;
        key[e] = valid ? KTr::key(tv[e]) : key_t(0);
        pay[e] = valid ? KTr::pay(tv[e]) : 0;
        dg[e] = valid ? digit_of(KTr::low32(key[e]), pp.shift0, mask) : 0xFFFFFFFFu;
      }
      scatter_tile<KB, BLOCK, E>(key, pay, dg, len, F2, pp.shift0, mask, s_stage, s_cnt, s_base, s_cur, s_delta,
                                 s_tmp, out, T.n);
#pragma unroll
      for (int e = 0; e < E; ++e) tv[e] = tn[e];
    }
  }
}

hipError_t launch_pass2(const Layout& L, int key_bytes, hipStream_t s) {
  dim3 grid(kNumXcd * kPass2WgPerXcd, L.ntables), block(kScatterBlock);
  if (key_bytes == 4) {
    const size_t lds = scatter_lds_bytes<4, kScatterTile4>(L.pp.F2);
    hipLaunchKernelGGL((k_pass2<4, kScatterBlock, kScatterTile4>), grid, block, lds, s, L.t[0], L.t[1], L.pp);
  } else {
    const size_t lds = scatter_lds_bytes<8, kScatterTile8>(L.pp.F2);
    hipLaunchKernelGGL((k_pass2<8, kScatterBlock, kScatterTile8>), grid, block, lds, s, L.t[0], L.t[1], L.pp);
  }
  return hipGetLastError();
}

}  // namespace pj
