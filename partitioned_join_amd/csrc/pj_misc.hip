// pj_misc.hip -- generators, device primitives and the phase-level kernels.
#include "pj_device.h"
#include "pj_internal.h"
#include "pj_misc.h"

namespace pj {

static int grid_for(int64_t n, int block, int cap = 8192) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return int(g);
}

#define PJ_GRID_STRIDE(i, n) \
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < (n); i += int64_t(gridDim.x) * blockDim.x)

// ---------------------------------------------------------------------------
// Generators (arithmetic identical to the host definitions in pj_device.h).
// ---------------------------------------------------------------------------
__global__ void k_gen_build_unique_i32(int32_t* keys, int64_t n, PermKey pk) {
  PJ_GRID_STRIDE(i, n) keys[i] = int32_t(perm_apply(uint64_t(i), pk));
}

__global__ void k_gen_build_dup_i32(int32_t* keys, int64_t n, int64_t m, PermKey pk, uint64_t seed) {
  PJ_GRID_STRIDE(i, n) keys[i] = int32_t(perm_apply(bounded(rand64(seed, uint64_t(i)), uint64_t(m)), pk));
}

__global__ void k_gen_probe_i32(int32_t* keys, int64_t n, int64_t domain, int64_t n_build, PermKey pk, uint64_t seed,
                                uint64_t thr, int64_t nm_base, int64_t nm_range) {
  PJ_GRID_STRIDE(j, n) {
    const uint64_t r1 = rand64(seed, 2 * uint64_t(j));
    const uint64_t r2 = rand64(seed, 2 * uint64_t(j) + 1);
    int64_t key;
    if ((r1 >> 32) < thr) key = int64_t(perm_apply(bounded(r2, uint64_t(n_build)), pk));
    else if (nm_range == 0) key = int64_t(perm_apply(uint64_t(n_build) + bounded(r2, uint64_t(domain - n_build)), pk));
    else key = nm_base + int64_t(bounded(r2, uint64_t(nm_range)));
    keys[j] = int32_t(key);
  }
}

__device__ __forceinline__ int64_t zipf_rank_dev(const double* lut, int64_t n, double u) {
  if (lut[0] >= u) return 0;
  int64_t lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (lut[mid] < u) lo = mid; else hi = mid;
  }
  return hi;
}

__global__ void k_gen_probe_zipf_i32(int32_t* keys, int64_t count, const double* lut, int64_t n, PermKey pk,
                                     uint64_t seed) {
  PJ_GRID_STRIDE(j, count) {
    const int64_t rank = zipf_rank_dev(lut, n, unit_double(rand64(seed, uint64_t(j))));
    keys[j] = int32_t(perm_apply(uint64_t(rank), pk));
  }
}

__global__ void k_gen_build_i64(int64_t* keys, int64_t n, uint64_t key_seed) {
  PJ_GRID_STRIDE(i, n) keys[i] = int64_t(rand64(key_seed, uint64_t(i)));
}

__global__ void k_gen_probe_i64(int64_t* keys, int64_t n, int64_t n_build, uint64_t key_seed, uint64_t seed,
                                uint64_t thr) {
  PJ_GRID_STRIDE(j, n) {
    const uint64_t r1 = rand64(seed, 2 * uint64_t(j));
    const uint64_t r2 = rand64(seed, 2 * uint64_t(j) + 1);
    const uint64_t idx = (r1 >> 32) < thr ? uint64_t(bounded(r2, uint64_t(n_build))) : uint64_t(n_build) + (r2 >> 2);
    keys[j] = int64_t(rand64(key_seed, idx));
  }
}

__global__ void k_gen_probe_zipf_i64(int64_t* keys, int64_t count, const double* lut, int64_t n_build,
                                     uint64_t key_seed, PermKey pk, uint64_t seed) {
  PJ_GRID_STRIDE(j, count) {
    const int64_t rank = zipf_rank_dev(lut, n_build, unit_double(rand64(seed, uint64_t(j))));
    keys[j] = int64_t(rand64(key_seed, perm_apply(uint64_t(rank), pk)));
  }
}

__global__ void k_row_ids(int32_t* p, int64_t n) {
  PJ_GRID_STRIDE(i, n) p[i] = int32_t(i);
}

hipError_t gen_build_unique_i32(int32_t* d, int64_t n, int64_t domain, uint64_t perm_seed, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_build_unique_i32, dim3(grid_for(n, 256)), dim3(256), 0, s, d, n,
                     perm_key(uint64_t(domain), perm_seed));
  return hipGetLastError();
}
hipError_t gen_build_dup_i32(int32_t* d, int64_t n, int64_t m, int64_t domain, uint64_t perm_seed, uint64_t seed,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_gen_build_dup_i32, dim3(grid_for(n, 256)), dim3(256), 0, s, d, n, m,
                     perm_key(uint64_t(domain), perm_seed), seed);
  return hipGetLastError();
}
hipError_t gen_probe_i32(int32_t* d, int64_t n, int64_t domain, int64_t n_build, uint64_t perm_seed, uint64_t seed,
                         double sel, int64_t nm_base, int64_t nm_range, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_probe_i32, dim3(grid_for(n, 256)), dim3(256), 0, s, d, n, domain, n_build,
                     perm_key(uint64_t(domain), perm_seed), seed, sel_threshold(sel), nm_base, nm_range);
  return hipGetLastError();
}
hipError_t gen_probe_zipf_i32(int32_t* d, int64_t count, const double* lut, int64_t n, uint64_t rank_perm_seed,
                              uint64_t seed, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_probe_zipf_i32, dim3(grid_for(count, 256)), dim3(256), 0, s, d, count, lut, n,
                     perm_key(uint64_t(n), rank_perm_seed), seed);
  return hipGetLastError();
}
hipError_t gen_build_i64(int64_t* d, int64_t n, uint64_t key_seed, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_build_i64, dim3(grid_for(n, 256)), dim3(256), 0, s, d, n, key_seed);
  return hipGetLastError();
}
hipError_t gen_probe_i64(int64_t* d, int64_t n, int64_t n_build, uint64_t key_seed, uint64_t seed, double sel,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_gen_probe_i64, dim3(grid_for(n, 256)), dim3(256), 0, s, d, n, n_build, key_seed, seed,
                     sel_threshold(sel));
  return hipGetLastError();
}
hipError_t gen_probe_zipf_i64(int64_t* d, int64_t count, const double* lut, int64_t n_build, uint64_t key_seed,
                              uint64_t rank_perm_seed, uint64_t seed, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_probe_zipf_i64, dim3(grid_for(count, 256)), dim3(256), 0, s, d, count, lut, n_build,
                     key_seed, perm_key(uint64_t(n_build), rank_perm_seed), seed);
  return hipGetLastError();
}
hipError_t gen_row_ids(int32_t* d, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_row_ids, dim3(grid_for(n, 256)), dim3(256), 0, s, d, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Primitives for parity tests.
// ---------------------------------------------------------------------------
__global__ void k_hash4(const uint32_t* k, int64_t n, uint32_t* out) { PJ_GRID_STRIDE(i, n) out[i] = murmur3(k[i]); }
__global__ void k_hash8(const uint64_t* k, int64_t n, uint32_t* out) { PJ_GRID_STRIDE(i, n) out[i] = murmur3(k[i]); }

__global__ void k_fastdiv(const int32_t* nv, int64_t count, FastDiv f, int32_t* q, int32_t* r) {
  PJ_GRID_STRIDE(i, count) {
    const int32_t x = nv[i];
    if (q) q[i] = fastdiv_div(x, f);
    if (r) r[i] = fastdiv_mod(x, f);
  }
}

hipError_t hash_keys(const void* keys, int64_t n, int key_bytes, uint32_t* out, hipStream_t s) {
  if (key_bytes == 4)
    hipLaunchKernelGGL(k_hash4, dim3(grid_for(n, 256)), dim3(256), 0, s, static_cast<const uint32_t*>(keys), n, out);
  else
    hipLaunchKernelGGL(k_hash8, dim3(grid_for(n, 256)), dim3(256), 0, s, static_cast<const uint64_t*>(keys), n, out);
  return hipGetLastError();
}

hipError_t fastdiv_run(const int32_t* nv, int64_t count, int32_t d, int32_t* q, int32_t* r, hipStream_t s) {
  hipLaunchKernelGGL(k_fastdiv, dim3(grid_for(count, 256)), dim3(256), 0, s, nv, count, fastdiv_make(d), q, r);
  return hipGetLastError();
}

template <typename K>
__global__ void k_project(const K* keys, const uint2* pairs, int64_t n, K* out) {
  PJ_GRID_STRIDE(i, n) {
    const uint2 p = pairs[i];
    out[2 * i] = keys[int32_t(p.x)];
    out[2 * i + 1] = K(int32_t(p.y));
  }
}

hipError_t project_reference(const void* keys, int key_bytes, const pj_pair* pairs, int64_t n, void* out,
                             hipStream_t s) {
  const uint2* pr = reinterpret_cast<const uint2*>(pairs);
  if (key_bytes == 4)
    hipLaunchKernelGGL(k_project<int32_t>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       static_cast<const int32_t*>(keys), pr, n, static_cast<int32_t*>(out));
  else
    hipLaunchKernelGGL(k_project<int64_t>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       static_cast<const int64_t*>(keys), pr, n, static_cast<int64_t*>(out));
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gather of payload columns through the join's pairs (cuDF's gather map):
// out_build[i] = build_col[pairs[i].build_idx], out_probe[i] =
// probe_col[pairs[i].probe_idx].  Each thread handles 4 pairs (two 16-byte
// pair loads); the column reads are random and the writes coalesced.
// ---------------------------------------------------------------------------
template <typename TB, typename TP>
__global__ __launch_bounds__(256) void k_gather(const uint2* pairs, int64_t n, const TB* bcol, const TP* pcol,
                                                TB* bout, TP* pout) {
  const int64_t groups = (n + 3) / 4;
  PJ_GRID_STRIDE(g, groups) {
    const int64_t i0 = g * 4;
    uint2 pr[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) pr[c] = i0 + c < n ? pairs[i0 + c] : make_uint2(0u, 0u);
    TB vb[4];
    TP vp[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (bcol && i0 + c < n) vb[c] = bcol[int32_t(pr[c].x)];
      if (pcol && i0 + c < n) vp[c] = pcol[int32_t(pr[c].y)];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (bcol && i0 + c < n) bout[i0 + c] = vb[c];
      if (pcol && i0 + c < n) pout[i0 + c] = vp[c];
    }
  }
}

template <typename TB, typename TP>
static hipError_t gather_t(const pj_pair* pairs, int64_t n, const void* bcol, const void* pcol, void* bout,
                           void* pout, hipStream_t s) {
  hipLaunchKernelGGL((k_gather<TB, TP>), dim3(grid_for((n + 3) / 4, 256)), dim3(256), 0, s,
                     reinterpret_cast<const uint2*>(pairs), n, static_cast<const TB*>(bcol),
                     static_cast<const TP*>(pcol), static_cast<TB*>(bout), static_cast<TP*>(pout));
  return hipGetLastError();
}

hipError_t gather_columns(const pj_pair* pairs, int64_t n, const void* bcol, int bbytes, const void* pcol,
                          int pbytes, void* bout, void* pout, hipStream_t s) {
  if (bbytes == 4 && pbytes == 4) return gather_t<int32_t, int32_t>(pairs, n, bcol, pcol, bout, pout, s);
  if (bbytes == 4 && pbytes == 8) return gather_t<int32_t, int64_t>(pairs, n, bcol, pcol, bout, pout, s);
  if (bbytes == 8 && pbytes == 4) return gather_t<int64_t, int32_t>(pairs, n, bcol, pcol, bout, pout, s);
  return gather_t<int64_t, int64_t>(pairs, n, bcol, pcol, bout, pout, s);
}

// ---------------------------------------------------------------------------
// Digit histogram (phase API): LDS bins, one global atomic per non-empty bin
// per workgroup.  d_hist must be zeroed.  tuple_stride: elements between keys
// (1 for SoA keys, 2 / 2 for 4-byte / 8-byte-key pj tuples).
// ---------------------------------------------------------------------------
template <typename K>
__global__ __launch_bounds__(1024) void k_digit_hist(const K* keys, int64_t n, int stride, int shift, int bits,
                                                    uint32_t* hist) {
  extern __shared__ uint32_t s_bins[];
  const int F = 1 << bits;
  const uint32_t mask = uint32_t(F - 1);
  for (int i = threadIdx.x; i < F; i += blockDim.x) s_bins[i] = 0;
  __syncthreads();
  PJ_GRID_STRIDE(i, n) {
    const uint32_t low = uint32_t(uint64_t(keys[i * stride]));
    atomicAdd(&s_bins[(low >> shift) & mask], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < F; i += blockDim.x)
    if (s_bins[i]) atomicAdd(&hist[i], s_bins[i]);
}

// Wide digits (more bins than fit in LDS): one global atomic per key.
template <typename K>
__global__ void k_digit_hist_global(const K* keys, int64_t n, int stride, int shift, int bits, uint32_t* hist) {
  const uint32_t mask = (1u << bits) - 1u;
  PJ_GRID_STRIDE(i, n) atomicAdd(&hist[(uint32_t(uint64_t(keys[i * stride])) >> shift) & mask], 1u);
}

hipError_t digit_hist(const void* keys, int64_t n, int key_bytes, int stride, int shift, int bits, uint32_t* hist,
                      hipStream_t s) {
  hipError_t e = launch_fill_u32(hist, 0, int64_t(1) << bits, s);
  if (e != hipSuccess || n == 0) return e;
  if (bits > 14) {
    if (key_bytes == 4)
      hipLaunchKernelGGL(k_digit_hist_global<uint32_t>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                         static_cast<const uint32_t*>(keys), n, stride, shift, bits, hist);
    else
      hipLaunchKernelGGL(k_digit_hist_global<uint64_t>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                         static_cast<const uint64_t*>(keys), n, stride, shift, bits, hist);
    return hipGetLastError();
  }
  const size_t lds = (size_t(1) << bits) * 4;
  const int g = grid_for(n, 1024 * 16, 1024);
  if (key_bytes == 4)
    hipLaunchKernelGGL(k_digit_hist<uint32_t>, dim3(g), dim3(1024), lds, s, static_cast<const uint32_t*>(keys), n,
                       stride, shift, bits, hist);
  else
    hipLaunchKernelGGL(k_digit_hist<uint64_t>, dim3(g), dim3(1024), lds, s, static_cast<const uint64_t*>(keys), n,
                       stride, shift, bits, hist);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Three-phase scan of uint32: block sums, scan of block sums (one workgroup),
// then per-block rescan with the carried offset.
// ---------------------------------------------------------------------------
constexpr int kScanBlock = 1024;
constexpr int kScanPer = 16;
constexpr int64_t kScanTile = int64_t(kScanBlock) * kScanPer;

__global__ __launch_bounds__(kScanBlock) void k_scan_sums(const uint32_t* in, int64_t n, uint32_t* sums) {
  __shared__ uint32_t s_tmp[kScanBlock / kWave + 1];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanPer;
  uint32_t v = 0;
  for (int k = 0; k < kScanPer; ++k)
    if (base + k < n) v += in[base + k];
  uint32_t total;
  block_exclusive_scan<kScanBlock>(v, s_tmp, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_top(uint32_t* sums, int64_t nb) {
  __shared__ uint32_t s_tmp[kScanBlock / kWave + 1];
  const int64_t per = (nb + kScanBlock - 1) / kScanBlock;
  const int64_t a = int64_t(threadIdx.x) * per, e = min(nb, a + per);
  uint32_t v = 0;
  for (int64_t i = a; i < e; ++i) v += sums[i];
  uint32_t total;
  uint32_t run = block_exclusive_scan<kScanBlock>(v, s_tmp, &total);
  for (int64_t i = a; i < e; ++i) {
    const uint32_t x = sums[i];
    sums[i] = run;
    run += x;
  }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const uint32_t* in, uint32_t* out, int64_t n,
                                                           const uint32_t* sums, int inclusive) {
  __shared__ uint32_t s_tmp[kScanBlock / kWave + 1];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanPer;
  uint32_t vals[kScanPer];
  uint32_t v = 0;
  for (int k = 0; k < kScanPer; ++k) {
    vals[k] = base + k < n ? in[base + k] : 0u;
    v += vals[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan<kScanBlock>(v, s_tmp, &total) + sums[blockIdx.x];
  for (int k = 0; k < kScanPer; ++k) {
    if (base + k < n) out[base + k] = inclusive ? run + vals[k] : run;
    run += vals[k];
  }
}

size_t scan_ws_bytes(int64_t n) { return size_t((n + kScanTile - 1) / kScanTile + 1) * 4 + 256; }

hipError_t scan_u32(const uint32_t* in, uint32_t* out, int64_t n, int inclusive, uint32_t* ws, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_sums, dim3(unsigned(nb)), dim3(kScanBlock), 0, s, in, n, ws);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanBlock), 0, s, ws, nb);
  hipLaunchKernelGGL(k_scan_apply, dim3(unsigned(nb)), dim3(kScanBlock), 0, s, in, out, n, ws, inclusive);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Block assignment (handle_skewed_data semantics).
// ---------------------------------------------------------------------------
__global__ void k_blocks_each(const uint32_t* off, int64_t total, int P, int base, uint32_t* blocks) {
  PJ_GRID_STRIDE(p, P) {
    const uint32_t end = (p == P - 1) ? uint32_t(total) : off[p + 1];
    const uint32_t size = end - off[p];
    blocks[p] = size == 0 ? 1u : (size - 1) / uint32_t(base) + 1u;
  }
}

__global__ void k_partition_index(const uint32_t* scanned, int P, uint32_t* pidx, int64_t max_blocks,
                                  uint32_t* d_total) {
  const uint32_t total = scanned[P - 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) *d_total = total;
  const int64_t lim = min(int64_t(total), max_blocks);
  PJ_GRID_STRIDE(b, lim) {
    // first p with scanned[p] > b
    int lo = -1, hi = P - 1;
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (scanned[mid] > uint32_t(b)) hi = mid; else lo = mid;
    }
    pidx[b] = uint32_t(hi);
  }
}

hipError_t assign_blocks(const uint32_t* off, int64_t total, int P, int base, uint32_t* scanned, uint32_t* pidx,
                         int64_t max_blocks, uint32_t* d_total, uint32_t* ws, hipStream_t s) {
  // scanned doubles as the per-partition block count before the scan
  hipLaunchKernelGGL(k_blocks_each, dim3(grid_for(P, 256)), dim3(256), 0, s, off, total, P, base, scanned);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = scan_u32(scanned, scanned, P, 1, ws, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_partition_index, dim3(256), dim3(256), 0, s, scanned, P, pidx, max_blocks, d_total);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pass-2 phase API set-up: treat the grouped input as the output of a single
// pass-1 segment (matc row 0 = 0, rows >= 1 = partition sizes; mat row 0 = 0).
// ---------------------------------------------------------------------------
__global__ void k_p2_prep(const uint32_t* off1, int F1, uint32_t nseg, uint32_t* matc, uint32_t* mat_row0, int P) {
  PJ_GRID_STRIDE(i, int64_t(nseg) * F1) {
    const int64_t s = i / F1, p = i % F1;
    matc[i] = s == 0 ? 0u : off1[p + 1] - off1[p];
  }
  PJ_GRID_STRIDE(q, P) mat_row0[q] = 0;
}

hipError_t p2_prep(const uint32_t* off1, int F1, uint32_t nseg, uint32_t* matc, uint32_t* mat_row0, int P,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_p2_prep, dim3(grid_for(int64_t(nseg) * F1 + P, 256)), dim3(256), 0, s, off1, F1, nseg, matc,
                     mat_row0, P);
  return hipGetLastError();
}

}  // namespace pj
