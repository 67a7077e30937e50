// pj_planner.hip -- learned join planner for gfx950.
//
//   k_sketch          one workgroup: 4096 probe keys sampled at fixed strides,
//                     a 1024-bin histogram of their MurmurHash3 values (sorted
//                     by count, so the features describe the skew profile and
//                     not which bins it falls in), six size/skew scalars, and
//                     the 64 most frequent sampled keys as heavy candidates.
//   k_count_cand      exact occurrence counts of the candidates in the whole
//                     probe column (LDS hash of the candidates, one pass;
//                     16-byte vector loads when the column is aligned).
//   k_gemm, ...       a 1030 -> 256 -> 256 -> 81 FP32 MLP: forward, backward
//                     and Adam.  The products are tiny (batch <= a few hundred)
//                     so one LDS-tiled FP32 GEMM with strided operands serves
//                     the forward and both backward products.
// The planner only chooses the plan; the join result never depends on it.
#include <cmath>

#include "pj_device.h"
#include "pj_internal.h"
#include "pj_planner.h"

namespace pj {

// ---------------------------------------------------------------------------
// Sketch
// ---------------------------------------------------------------------------
// Bitonic sort (ascending) of PER * 1024 uint64 held PER per thread by a
// 1024-thread workgroup, element i = PER * tid + c: a compare-exchange partner
// i ^ j lies in the same thread for j < PER (register swaps), in another lane
// of the same wave for j < PER * 64 (shuffles), and in another wave otherwise
// (through LDS, two barriers).  Only the stages whose partners cross waves
// synchronise the workgroup (10 of 78 for 4096 values).  Ends with the
// sorted values in v and in s[0 .. PER * 1024).
template <int PER>
__device__ __forceinline__ void bitonic_sort_block(uint64_t (&v)[PER], uint64_t* s) {
  constexpr uint32_t N = uint32_t(PER) * 1024u;
  const uint32_t tid = threadIdx.x;
  for (uint32_t size = 2; size <= N; size <<= 1) {
    for (uint32_t j = size >> 1; j > 0; j >>= 1) {
      uint64_t w[PER];
      if (j >= uint32_t(PER) * kWave) {
#pragma unroll
        for (int c = 0; c < PER; ++c) s[PER * tid + c] = v[c];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < PER; ++c) w[c] = s[(PER * tid + c) ^ j];
        __syncthreads();
      } else if (j >= uint32_t(PER)) {
#pragma unroll
        for (int c = 0; c < PER; ++c) w[c] = __shfl_xor(v[c], int(j / PER), kWave);
      } else {
#pragma unroll
        for (int c = 0; c < PER; ++c) w[c] = v[c ^ j];
      }
#pragma unroll
      for (int c = 0; c < PER; ++c) {
        const uint32_t i = PER * tid + c;
        const bool keep_min = ((i & j) == 0) == ((i & size) == 0);
        const uint64_t a = v[c], b = w[c];
        v[c] = keep_min ? (a < b ? a : b) : (a < b ? b : a);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < PER; ++c) s[PER * tid + c] = v[c];
  __syncthreads();
}

template <int KB>
__global__ __launch_bounds__(1024) void k_sketch(const void* keys, int64_t n, int64_t n_build, float* feat,
                                                 uint64_t* cand_keys, uint32_t* cand_counts) {
  using key_t = typename KeyT<KB>::key_t;
  __shared__ uint64_t s_keys[kPlanSample];
  __shared__ uint64_t s_sel[kPlanSample];
  __shared__ uint32_t s_hist[kPlanBins];
  __shared__ float s_red[2 * 1024 / kWave];
  const int tid = threadIdx.x;
  const key_t* k = static_cast<const key_t*>(keys);
  static_assert(kPlanSample == 4 * 1024 && kPlanBins == 1024, "sort shapes");
  for (int b = tid; b < kPlanBins; b += 1024) s_hist[b] = 0;
  __syncthreads();
  uint64_t v[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int j = 4 * tid + c;
    v[c] = 0;
    if (n > 0) {
      const int64_t pos = ((2 * int64_t(j) + 1) * n) / (2 * int64_t(kPlanSample));  // < n
      const key_t key = k[pos];
      v[c] = uint64_t(key);
      if (KB == 4) v[c] &= 0xffffffffull;
      atomicAdd(&s_hist[murmur3(key) & uint32_t(kPlanBins - 1)], 1u);
    }
  }
  bitonic_sort_block<4>(v, s_keys);
  // Runs of equal keys: a run's first position holds its length (found by a
  // binary search for the run's end).  The sort key (4096 - count) << 12 |
  // position orders runs by count, then by key.
  uint64_t sel[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int i = 4 * tid + c;
    uint64_t sk = (uint64_t(kPlanSample) << 12) | uint64_t(i);
    if (n > 0 && (i == 0 || s_keys[i] != s_keys[i - 1])) {
      int lo = i, hi = kPlanSample;  // s_keys[lo] == key, the end is in (lo, hi]
      const uint64_t key = s_keys[i];
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_keys[mid] == key) lo = mid; else hi = mid;
      }
      sk = (uint64_t(kPlanSample - (hi - i)) << 12) | uint64_t(i);
    }
    sel[c] = sk;
  }
  // Sample collision probability: sum over runs of (count / 4096)^2.
  float sq = 0.f;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const uint32_t cnt = uint32_t(kPlanSample) - uint32_t(sel[c] >> 12);
    sq += float(cnt) * float(cnt);
  }
  for (int d = kWave / 2; d > 0; d >>= 1) sq += __shfl_xor(sq, d, kWave);
  if ((tid & (kWave - 1)) == 0) s_red[tid / kWave] = sq;
  bitonic_sort_block<4>(sel, s_sel);
  if (tid < kPlanCand) {
    const uint64_t sk = s_sel[tid];
    const uint32_t c = uint32_t(kPlanSample) - uint32_t(sk >> 12);
    cand_keys[tid] = c ? s_keys[sk & 0xfffu] : 0ull;
    cand_counts[tid] = c;
  }
  float top = 0.f;
  if (tid == 0) top = float(uint32_t(kPlanSample) - uint32_t(s_sel[0] >> 12));
  __syncthreads();
  // Histogram bins sorted by count, descending (reuse s_keys).
  uint64_t hb[1] = {~uint64_t(s_hist[tid])};
  bitonic_sort_block<1>(hb, s_keys);
  feat[tid] = float(~hb[0]) / float(kPlanSample);
  if (tid == 0) {
    float tot = 0.f;
    for (int w = 0; w < 1024 / kWave; ++w) tot += s_red[w];
    const double nb = double(n_build) + 1.0, ns = double(n) + 1.0;
    feat[kPlanBins + 0] = float(log2(nb) / 32.0);
    feat[kPlanBins + 1] = float(log2(ns) / 32.0);
    feat[kPlanBins + 2] = float(log2(ns / nb) / 8.0);
    feat[kPlanBins + 3] = float(KB) / 8.0f;
    feat[kPlanBins + 4] = n > 0 ? top / float(kPlanSample) : 0.f;
    feat[kPlanBins + 5] = n > 0 ? tot / (float(kPlanSample) * float(kPlanSample)) : 0.f;
  }
}

// ---------------------------------------------------------------------------
// Exact candidate counts: the valid candidates (sample count > 0) sit in a
// 128-slot LDS open-addressing table; every probe key is looked up once.  The
// table's slot is the top bits of a multiplicative hash (one multiply per
// key; the counts do not depend on it).
// ---------------------------------------------------------------------------
template <int SLOTS>
__device__ __forceinline__ uint32_t cand_slot(uint32_t k) {
  return (k * 0x9E3779B1u) >> (32 - __builtin_ctz(uint32_t(SLOTS)));
}
template <int SLOTS>
__device__ __forceinline__ uint32_t cand_slot(uint64_t k) {
  return ((uint32_t(k) ^ (uint32_t(k >> 32) * 0x85EBCA77u)) * 0x9E3779B1u) >> (32 - __builtin_ctz(uint32_t(SLOTS)));
}
template <int KB>
__global__ __launch_bounds__(256) void k_count_cand(const void* keys, int64_t n, const uint64_t* cand_keys,
                                                    const uint32_t* cand_sample_counts,
                                                    unsigned long long* counts) {
  using key_t = typename KeyT<KB>::key_t;
  constexpr int SLOTS = 2 * kPlanCand;
  __shared__ uint64_t s_key[SLOTS];
  __shared__ int32_t s_idx[SLOTS];
  __shared__ uint32_t s_cnt[kPlanCand];
  const int tid = threadIdx.x;
  for (int i = tid; i < SLOTS; i += 256) s_idx[i] = -1;
  for (int i = tid; i < kPlanCand; i += 256) s_cnt[i] = 0;
  __syncthreads();
  if (tid == 0) {
    for (int c = 0; c < kPlanCand; ++c) {
      if (cand_sample_counts[c] == 0) continue;
      const key_t key = key_t(cand_keys[c]);
      uint32_t s = cand_slot<SLOTS>(key);
      while (s_idx[s] >= 0) s = (s + 1) & uint32_t(SLOTS - 1);
      s_key[s] = uint64_t(key);
      s_idx[s] = c;
    }
  }
  __syncthreads();
  const key_t* k = static_cast<const key_t*>(keys);
  for (int64_t i = blockIdx.x * int64_t(256) + tid; i < n; i += int64_t(gridDim.x) * 256) {
    const key_t key = k[i];
    uint64_t v = uint64_t(key);
    if (KB == 4) v &= 0xffffffffull;
    uint32_t s = cand_slot<SLOTS>(key);
    while (s_idx[s] >= 0) {  // at most 64 of the 128 slots are used
      if (s_key[s] == v) {
        atomicAdd(&s_cnt[s_idx[s]], 1u);
        break;
      }
      s = (s + 1) & uint32_t(SLOTS - 1);
    }
  }
  __syncthreads();
  for (int c = tid; c < kPlanCand; c += 256)
    if (s_cnt[c]) atomicAdd(&counts[c], (unsigned long long)s_cnt[c]);
}

// The same count over a 16-byte-aligned column, read as 16-byte vectors with
// two loads in flight per thread; the n % (16 / key bytes) last keys are read
// one by one by the first workgroup.  The table has 16 slots per candidate,
// so a key that is no candidate (nearly all of them) mostly ends its walk at
// the first slot it reads.
template <int KB>
__global__ __launch_bounds__(256) void k_count_cand_vec(const void* keys, int64_t n, const uint64_t* cand_keys,
                                                        const uint32_t* cand_sample_counts,
                                                        unsigned long long* counts) {
  using key_t = typename KeyT<KB>::key_t;
  constexpr int SLOTS = 16 * kPlanCand;
  constexpr int PER = 16 / KB;  // keys per vector
  __shared__ uint64_t s_key[SLOTS];
  __shared__ int32_t s_idx[SLOTS];
  __shared__ uint32_t s_cnt[kPlanCand];
  const int tid = threadIdx.x;
  for (int i = tid; i < SLOTS; i += 256) s_idx[i] = -1;
  for (int i = tid; i < kPlanCand; i += 256) s_cnt[i] = 0;
  __syncthreads();
  if (tid == 0) {
    for (int c = 0; c < kPlanCand; ++c) {
      if (cand_sample_counts[c] == 0) continue;
      const key_t key = key_t(cand_keys[c]);
      uint32_t s = cand_slot<SLOTS>(key);
      while (s_idx[s] >= 0) s = (s + 1) & uint32_t(SLOTS - 1);
      s_key[s] = uint64_t(key);
      s_idx[s] = c;
    }
  }
  __syncthreads();
  auto lookup = [&](key_t key) {
    uint64_t v = uint64_t(key);
    if (KB == 4) v &= 0xffffffffull;
    uint32_t s = cand_slot<SLOTS>(key);
    while (s_idx[s] >= 0) {  // at most 64 of the slots are used
      if (s_key[s] == v) {
        atomicAdd(&s_cnt[s_idx[s]], 1u);
        break;
      }
      s = (s + 1) & uint32_t(SLOTS - 1);
    }
  };
  auto lookup_vec = [&](const uint4 q) {
    if constexpr (KB == 4) {
      lookup(q.x);
      lookup(q.y);
      lookup(q.z);
      lookup(q.w);
    } else {
      lookup(uint64_t(q.x) | (uint64_t(q.y) << 32));
      lookup(uint64_t(q.z) | (uint64_t(q.w) << 32));
    }
  };
  const uint4* k4 = static_cast<const uint4*>(keys);
  const int64_t nv = n / PER, stride = int64_t(gridDim.x) * 256;
  int64_t i = blockIdx.x * int64_t(256) + tid;
  for (; i + stride < nv; i += 2 * stride) {
    const uint4 a = k4[i], b = k4[i + stride];
    lookup_vec(a);
    lookup_vec(b);
  }
  if (i < nv) lookup_vec(k4[i]);
  if (blockIdx.x == 0 && int64_t(tid) < n - nv * PER) lookup(static_cast<const key_t*>(keys)[nv * PER + tid]);
  __syncthreads();
  for (int c = tid; c < kPlanCand; c += 256)
    if (s_cnt[c]) atomicAdd(&counts[c], (unsigned long long)s_cnt[c]);
}

hipError_t planner_sketch(const void* keys, int64_t n, int key_bytes, int64_t n_build, float* feat,
                          uint64_t* cand_keys, uint32_t* cand_counts, hipStream_t s) {
  if (key_bytes == 4)
    hipLaunchKernelGGL(k_sketch<4>, dim3(1), dim3(1024), 0, s, keys, n, n_build, feat, cand_keys, cand_counts);
  else
    hipLaunchKernelGGL(k_sketch<8>, dim3(1), dim3(1024), 0, s, keys, n, n_build, feat, cand_keys, cand_counts);
  return hipGetLastError();
}

hipError_t planner_count_candidates(const void* keys, int64_t n, int key_bytes, const uint64_t* cand_keys,
                                    const uint32_t* cand_sample_counts, unsigned long long* counts,
                                    hipStream_t s) {
  hipError_t e = launch_fill_u64(reinterpret_cast<uint64_t*>(counts), 0, kPlanCand, s);
  if (e != hipSuccess || n == 0) return e;
  if (reinterpret_cast<uintptr_t>(keys) % 16 == 0) {
    int64_t g = (n / (16 / key_bytes) + 511) / 512;
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    if (key_bytes == 4)
      hipLaunchKernelGGL(k_count_cand_vec<4>, dim3(int(g)), dim3(256), 0, s, keys, n, cand_keys, cand_sample_counts,
                         counts);
    else
      hipLaunchKernelGGL(k_count_cand_vec<8>, dim3(int(g)), dim3(256), 0, s, keys, n, cand_keys, cand_sample_counts,
                         counts);
    return hipGetLastError();
  }
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  if (key_bytes == 4)
    hipLaunchKernelGGL(k_count_cand<4>, dim3(int(g)), dim3(256), 0, s, keys, n, cand_keys, cand_sample_counts,
                       counts);
  else
    hipLaunchKernelGGL(k_count_cand<8>, dim3(int(g)), dim3(256), 0, s, keys, n, cand_keys, cand_sample_counts,
                       counts);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// MLP kernels
// ---------------------------------------------------------------------------
// C[M x N] = act((accum ? C : 0) + op(A) op(B) + bias), where op(A)(m, k) =
// A[m * sam + k * sak] and op(B)(k, n) = B[k * sbk + n * sbn].  64 x 64 tiles,
// 256 threads with 4 x 4 outputs each, K staged through LDS 16 at a time.
__global__ __launch_bounds__(256) void k_gemm(const float* __restrict__ A, int64_t sam, int64_t sak,
                                              const float* __restrict__ B, int64_t sbk, int64_t sbn, float* C, int M,
                                              int N, int K, const float* bias, int relu, int accum) {
  constexpr int TM = 64, TN = 64, TK = 16;
  __shared__ float As[TK][TM + 4];
  __shared__ float Bs[TK][TN + 4];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  float acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
  for (int k0 = 0; k0 < K; k0 += TK) {
    for (int e = threadIdx.x; e < TM * TK; e += 256) {
      const int mm = e / TK, kk = e % TK, m = m0 + mm, k = k0 + kk;
      As[kk][mm] = (m < M && k < K) ? A[int64_t(m) * sam + int64_t(k) * sak] : 0.f;
    }
    for (int e = threadIdx.x; e < TK * TN; e += 256) {
      const int kk = e / TN, nn = e % TN, k = k0 + kk, n = n0 + nn;
      Bs[kk][nn] = (k < K && n < N) ? B[int64_t(k) * sbk + int64_t(n) * sbn] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = m0 + ty * 4 + i, n = n0 + tx * 4 + j;
      if (m < M && n < N) {
        float v = acc[i][j];
        if (accum) v += C[int64_t(m) * N + n];
        if (bias) v += bias[n];
        if (relu) v = v > 0.f ? v : 0.f;
        C[int64_t(m) * N + n] = v;
      }
    }
}

static hipError_t gemm(const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk, int64_t sbn, float* C,
                       int M, int N, int K, const float* bias, bool relu, hipStream_t s) {
  dim3 grid((N + 63) / 64, (M + 63) / 64);
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, A, sam, sak, B, sbk, sbn, C, M, N, K, bias, relu ? 1 : 0, 0);
  return hipGetLastError();
}

// db[n] = sum over rows of dY[r][n].
__global__ void k_colsum(const float* dY, int rows, int N, float* db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float acc = 0.f;
  for (int r = 0; r < rows; ++r) acc += dY[int64_t(r) * N + n];
  db[n] = acc;
}

// dH[i] = 0 where the ReLU output H[i] is 0.
__global__ void k_relu_mask(float* dH, const float* H, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    if (!(H[i] > 0.f)) dH[i] = 0.f;
}

// Softmax cross-entropy on the bits-1, bits-2 and pass heads and binary
// cross-entropy on the 64 heavy flags; labels[b] = {bits1 class, bits2 class
// or -1 (one pass), pass class}, heavy[b][j] = 1, 0 or -1 (no candidate).
// Writes dOut (mean over the batch) and the mean loss.  One workgroup.
__global__ __launch_bounds__(256) void k_heads_loss(const float* out, const int32_t* labels, const float* heavy,
                                                    int batch, float* dout, float* loss) {
  __shared__ float s_red[256 / kWave];
  float lsum = 0.f;
  const float inv = 1.f / float(batch);
  for (int b = threadIdx.x; b < batch; b += 256) {
    const float* o = out + int64_t(b) * kPlanOut;
    float* g = dout + int64_t(b) * kPlanOut;
    auto ce = [&](int off, int ncls, int label) {
      if (label < 0) {
        for (int c = 0; c < ncls; ++c) g[off + c] = 0.f;
        return 0.f;
      }
      float mx = o[off];
      for (int c = 1; c < ncls; ++c) mx = fmaxf(mx, o[off + c]);
      float z = 0.f;
      for (int c = 0; c < ncls; ++c) z += expf(o[off + c] - mx);
      for (int c = 0; c < ncls; ++c) g[off + c] = (expf(o[off + c] - mx) / z - (c == label ? 1.f : 0.f)) * inv;
      return logf(z) + mx - o[off + label];  // o and g are different buffers
    };
    const int32_t* lb = labels + int64_t(b) * 3;
    float l = ce(0, kPlanBitsCls, lb[0]);
    l += ce(kPlanBitsCls, kPlanBitsCls, lb[1]);
    l += ce(2 * kPlanBitsCls, kPlanPassCls, lb[2]);
    const int hoff = 2 * kPlanBitsCls + kPlanPassCls;
    for (int j = 0; j < kPlanCand; ++j) {
      const float t = heavy[int64_t(b) * kPlanCand + j];
      const float x = o[hoff + j];
      if (t < 0.f) {
        g[hoff + j] = 0.f;
        continue;
      }
      const float p = 1.f / (1.f + expf(-x));
      g[hoff + j] = (p - t) * inv / float(kPlanCand);
      // log(1 + e^-|x|) + max(x, 0) - t x, written to stay finite
      l += (fmaxf(x, 0.f) - t * x + log1pf(expf(-fabsf(x)))) / float(kPlanCand);
    }
    lsum += l;
  }
  for (int d = kWave / 2; d > 0; d >>= 1) lsum += __shfl_xor(lsum, d, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) s_red[threadIdx.x / kWave] = lsum;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < 256 / kWave; ++w) t += s_red[w];
    *loss = t * inv;
  }
}

__global__ void k_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float bc1, float bc2) {
  constexpr float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  }
}

// Scratch (floats): H1, H2, dH2, dH1 [batch x H]; logits and dOut
// [batch x O]; gradients [params].
size_t planner_scratch_floats(int batch) {
  return 4 * size_t(batch) * kPlanHidden + 2 * size_t(batch) * kPlanOut + kPlanParams + 64;
}

static hipError_t forward(const float* w, const float* x, int batch, float* h1, float* h2, float* out,
                          hipStream_t s) {
  const int F = kPlanFeat, H = kPlanHidden, O = kPlanOut;
  hipError_t e = gemm(x, F, 1, w + kPlanW1, H, 1, h1, batch, H, F, w + kPlanB1, true, s);
  if (e == hipSuccess) e = gemm(h1, H, 1, w + kPlanW2, H, 1, h2, batch, H, H, w + kPlanB2, true, s);
  if (e == hipSuccess) e = gemm(h2, H, 1, w + kPlanW3, O, 1, out, batch, O, H, w + kPlanB3, false, s);
  return e;
}

// One sample (planning a join): matrix-vector products split over K, so
// every layer reads its weights with many workgroups.  Workgroup w of a layer
// takes input rows [w * KC, (w + 1) * KC) and writes partial sums part[w][n];
// the next layer first adds its input's partials, the bias and the ReLU.
// Fixed split and summation order: the logits are deterministic.
constexpr int kGemvBlock = 256;
constexpr int kGemvSplit1 = 16, kGemvSplit2 = 8, kGemvSplit3 = 8;

__global__ __launch_bounds__(kGemvBlock) void k_gemv(const float* x, const float* x_parts, int x_nparts,
                                                     const float* x_bias, const float* W, int K, int N, int KC,
                                                     float* out_parts) {
  __shared__ float s_x[128];
  const int k0 = blockIdx.x * KC, k1 = min(K, k0 + KC);
  for (int k = k0 + threadIdx.x; k < k1; k += kGemvBlock) {
    float v;
    if (x_parts) {  // hidden layer input: sum of partials + bias, ReLU
      v = x_bias[k];
      for (int p = 0; p < x_nparts; ++p) v += x_parts[size_t(p) * K + k];
      v = v > 0.f ? v : 0.f;
    } else {
      v = x[k];
    }
    s_x[k - k0] = v;
  }
  __syncthreads();
  for (int n = threadIdx.x; n < N; n += kGemvBlock) {
    float acc = 0.f;
    for (int k = k0; k < k1; ++k) acc = fmaf(s_x[k - k0], W[size_t(k) * N + n], acc);
    out_parts[size_t(blockIdx.x) * N + n] = acc;
  }
}

__global__ __launch_bounds__(kGemvBlock) void k_gemv_final(const float* parts, int nparts, const float* bias, int N,
                                                           float* out) {
  for (int n = threadIdx.x; n < N; n += kGemvBlock) {
    float v = bias[n];
    for (int p = 0; p < nparts; ++p) v += parts[size_t(p) * N + n];
    out[n] = v;
  }
}

static hipError_t forward_one(const float* w, const float* x, float* out, float* scratch, hipStream_t s) {
  const int F = kPlanFeat, H = kPlanHidden, O = kPlanOut;
  float* p1 = scratch;                                  // [kGemvSplit1][H]
  float* p2 = p1 + size_t(kGemvSplit1) * H;             // [kGemvSplit2][H]
  float* p3 = p2 + size_t(kGemvSplit2) * H;             // [kGemvSplit3][O]
  const int kc1 = (F + kGemvSplit1 - 1) / kGemvSplit1, kc2 = H / kGemvSplit2, kc3 = H / kGemvSplit3;
  hipLaunchKernelGGL(k_gemv, dim3(kGemvSplit1), dim3(kGemvBlock), 0, s, x, nullptr, 0, nullptr, w + kPlanW1, F, H,
                     kc1, p1);
  hipLaunchKernelGGL(k_gemv, dim3(kGemvSplit2), dim3(kGemvBlock), 0, s, nullptr, p1, kGemvSplit1, w + kPlanB1,
                     w + kPlanW2, H, H, kc2, p2);
  hipLaunchKernelGGL(k_gemv, dim3(kGemvSplit3), dim3(kGemvBlock), 0, s, nullptr, p2, kGemvSplit2, w + kPlanB2,
                     w + kPlanW3, H, O, kc3, p3);
  hipLaunchKernelGGL(k_gemv_final, dim3(1), dim3(kGemvBlock), 0, s, p3, kGemvSplit3, w + kPlanB3, O, out);
  return hipGetLastError();
}
static_assert((kPlanFeat + kGemvSplit1 - 1) / kGemvSplit1 <= 128 && kPlanHidden / kGemvSplit2 <= 128,
              "a workgroup's input rows fit its LDS slice");

hipError_t planner_forward(const float* w, const float* x, int batch, float* out, float* scratch, hipStream_t s) {
  if (batch == 1) return forward_one(w, x, out, scratch, s);
  float* h1 = scratch;
  float* h2 = h1 + size_t(batch) * kPlanHidden;
  return forward(w, x, batch, h1, h2, out, s);
}

hipError_t planner_train_step(float* w, float* m, float* v, int step, float lr, const float* x, const int32_t* labels,
                              const float* heavy, int batch, float* loss, float* scratch, hipStream_t s) {
  const int F = kPlanFeat, H = kPlanHidden, O = kPlanOut;
  float* h1 = scratch;
  float* h2 = h1 + size_t(batch) * H;
  float* dh2 = h2 + size_t(batch) * H;
  float* dh1 = dh2 + size_t(batch) * H;
  float* out = dh1 + size_t(batch) * H;
  float* g_out = out + size_t(batch) * O;
  float* grad = g_out + size_t(batch) * O;
  hipError_t e = forward(w, x, batch, h1, h2, out, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_heads_loss, dim3(1), dim3(256), 0, s, out, labels, heavy, batch, g_out, loss);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // layer 3: dW3 = H2^T dOut, db3 = colsum(dOut), dH2 = dOut W3^T (masked by H2 > 0)
  if ((e = gemm(h2, 1, H, g_out, O, 1, grad + kPlanW3, H, O, batch, nullptr, false, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_colsum, dim3((O + 255) / 256), dim3(256), 0, s, g_out, batch, O, grad + kPlanB3);
  if ((e = gemm(g_out, O, 1, w + kPlanW3, 1, O, dh2, batch, H, O, nullptr, false, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_relu_mask, dim3(64), dim3(256), 0, s, dh2, h2, int64_t(batch) * H);
  // layer 2
  if ((e = gemm(h1, 1, H, dh2, H, 1, grad + kPlanW2, H, H, batch, nullptr, false, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_colsum, dim3((H + 255) / 256), dim3(256), 0, s, dh2, batch, H, grad + kPlanB2);
  if ((e = gemm(dh2, H, 1, w + kPlanW2, 1, H, dh1, batch, H, H, nullptr, false, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_relu_mask, dim3(64), dim3(256), 0, s, dh1, h1, int64_t(batch) * H);
  // layer 1
  if ((e = gemm(x, 1, F, dh1, H, 1, grad + kPlanW1, F, H, batch, nullptr, false, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_colsum, dim3((H + 255) / 256), dim3(256), 0, s, dh1, batch, H, grad + kPlanB1);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const double bc1 = 1.0 - std::pow(0.9, double(step)), bc2 = 1.0 - std::pow(0.999, double(step));
  hipLaunchKernelGGL(k_adam, dim3(512), dim3(256), 0, s, w, grad, m, v, int64_t(kPlanParams), lr, float(bc1),
                     float(bc2));
  return hipGetLastError();
}

}  // namespace pj
