// pj_join.hip -- per-partition LDS hash join for gfx950.
//
//   k_join_plan  work list: partition q with |R_q| build and |S_q| probe tuples
//                gives ceil(|R_q|/base) x ceil(|S_q|/slice) items (build chunk x
//                probe slice), so heavy partitions are split on both sides.
//                Device-side, no host round trip (the reference's
//                handle_skewed_data reads the block count back to the host,
//                src/skewed_data_handler.cuh:83, and splits only the build side).
//   k_join       persistent workgroups; per item: LDS open-addressing table of
//                2*|chunk| slots (occupancy <= 0.5), MurmurHash3 seed 0, slot =
//                |(int)h % size| with int_fastdiv, linear probing with LDS CAS
//                (src/hash_join_kernels.cuh:84-129), then the probe slice walks
//                the table (:134-181).  Matches are staged in LDS and flushed
//                into output chunks of kOutChunk pairs that each workgroup claims
//                with one global atomic (a same-address device atomic costs
//                ~12 ns on MI355X, so claims stay rare).  The empty-slot
//                sentinel is a key whose low bits differ from the partition
//                digit, so every real key value can be joined (the reference
//                reserves -1, :20,86,106).
//   k_defrag_*   the last chunk of each workgroup is partly filled; pairs that
//                lie past the total are moved into those holes so the output
//                is dense in [0, total).
#include "pj_device.h"
#include "pj_internal.h"

namespace pj {

int join_table_capacity(int key_bytes) { return key_bytes == 4 ? 7936 : 5120; }

static size_t join_lds_bytes(int key_bytes, int table_slots) {
  return size_t(table_slots) * (key_bytes == 4 ? 8 : 12) + size_t(kJoinStage) * 8 + 256;
}

// ---------------------------------------------------------------------------
// k_join_plan: item_scan[q] = exclusive scan of items per partition;
// item_scan[P] = total.  Also resets the output chunk counter.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_join_plan(const uint32_t* roff, const uint32_t* soff, int P, int base,
                                                    int64_t slice, uint64_t* item_scan, uint32_t* chunk_ctr) {
  __shared__ uint64_t s_tmp[1024 / kWave + 1];
  const int per = (P + 1023) / 1024;
  const int a = threadIdx.x * per, e = min(P, a + per);
  uint64_t sum = 0;
  for (int q = a; q < e; ++q) {
    const uint64_t nr = roff[q + 1] - roff[q], ns = soff[q + 1] - soff[q];
    if (nr && ns) sum += ((nr + base - 1) / base) * ((ns + slice - 1) / slice);
  }
  uint64_t total;
  uint64_t run = block_exclusive_scan64<1024>(sum, s_tmp, &total);
  for (int q = a; q < e; ++q) {
    const uint64_t nr = roff[q + 1] - roff[q], ns = soff[q + 1] - soff[q];
    item_scan[q] = run;
    if (nr && ns) run += ((nr + base - 1) / base) * ((ns + slice - 1) / slice);
  }
  if (threadIdx.x == 0) {
    item_scan[P] = total;
    *chunk_ctr = 0;
  }
}

// ---------------------------------------------------------------------------
// Output staging shared by the join workgroup.
// ---------------------------------------------------------------------------
struct StageCtl {
  uint32_t n;       // staged pairs
  uint32_t chunk;   // current output chunk
  uint32_t fill;    // pairs written to the current chunk
  uint32_t has_chunk;
  uint32_t a;       // pairs of this flush that go to the current chunk
  uint32_t pad;
  uint64_t dst0, dst1;
};

__device__ __forceinline__ void store_pair(const JoinArgs& A, uint64_t v, uint2 pr, uint64_t spill_cap) {
  if (v < uint64_t(A.capacity)) {
    reinterpret_cast<uint2*>(A.out)[v] = pr;
  } else if (v - uint64_t(A.capacity) < spill_cap) {
    reinterpret_cast<uint2*>(A.spill)[v - uint64_t(A.capacity)] = pr;
  }
}

// Called by all threads after a barrier (ctl->n stable).  Ends with a barrier.
template <int BLOCK>
__device__ __forceinline__ void flush_stage(const JoinArgs& A, StageCtl* ctl, const uint2* s_stage) {
  const uint64_t spill_cap = uint64_t(gridDim.x) * uint64_t(kOutChunk);
  if (threadIdx.x == 0) {
    const uint32_t n = ctl->n;
    const uint32_t avail = ctl->has_chunk ? uint32_t(kOutChunk) - ctl->fill : 0u;
    const uint32_t a = n < avail ? n : avail;
    ctl->dst0 = uint64_t(ctl->chunk) * uint64_t(kOutChunk) + ctl->fill;
    ctl->a = a;
    if (n > a) {
      const uint32_t c2 = atomicAdd(A.chunk_ctr, 1u);
      ctl->dst1 = uint64_t(c2) * uint64_t(kOutChunk);
      ctl->chunk = c2;
      ctl->fill = n - a;
      ctl->has_chunk = 1;
    } else {
      ctl->fill += a;
    }
  }
  __syncthreads();
  const uint32_t n = ctl->n, a = ctl->a;
  const uint64_t d0 = ctl->dst0, d1 = ctl->dst1;
  for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
    const uint64_t v = i < a ? d0 + i : d1 + (i - a);
    store_pair(A, v, s_stage[i], spill_cap);
  }
  __syncthreads();
  if (threadIdx.x == 0) ctl->n = 0;
  __syncthreads();
}

__device__ __forceinline__ uint32_t lds_cas(uint32_t* p, uint32_t cmp, uint32_t val) { return atomicCAS(p, cmp, val); }
__device__ __forceinline__ uint64_t lds_cas(uint64_t* p, uint64_t cmp, uint64_t val) {
  return uint64_t(atomicCAS(reinterpret_cast<unsigned long long*>(p), (unsigned long long)cmp,
                            (unsigned long long)val));
}

// ---------------------------------------------------------------------------
// k_join
// ---------------------------------------------------------------------------
template <int KB, int BLOCK, bool UNIQUE>
__global__ __launch_bounds__(BLOCK) void k_join(JoinArgs A) {
  using KTr = KeyT<KB>;
  using key_t = typename KTr::key_t;
  using tup_t = typename KTr::tup_t;
  extern __shared__ __align__(16) uint8_t s_raw[];
  uint2* s_stage = reinterpret_cast<uint2*>(s_raw);
  key_t* s_tkeys = reinterpret_cast<key_t*>(s_raw + size_t(kJoinStage) * 8);
  int32_t* s_tpays = reinterpret_cast<int32_t*>(s_tkeys + A.table_cap);
  __shared__ StageCtl ctl;
  __shared__ uint32_t s_rlo, s_rhi, s_slo, s_shi, s_q;
  __shared__ FastDiv s_fd;

  const int tid = threadIdx.x;
  const uint32_t lane = lane_id();
  const tup_t* rt = static_cast<const tup_t*>(A.rt);
  const tup_t* st = static_cast<const tup_t*>(A.st);
  const uint64_t total_items = A.item_scan[A.P];
  if (tid == 0) {
    ctl.n = 0;
    ctl.fill = 0;
    ctl.chunk = 0;
    ctl.has_chunk = 0;
  }
  __syncthreads();

  for (uint64_t item = blockIdx.x; item < total_items; item += gridDim.x) {
    if (tid == 0) {
      int lo = 0, hi = A.P;  // item_scan[lo] <= item < item_scan[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (A.item_scan[mid] <= item) lo = mid; else hi = mid;
      }
      const int q = lo;
      const uint64_t local = item - A.item_scan[q];
      const uint32_t r0 = A.roff[q], r1 = min(A.roff[q + 1], A.nr), s0 = A.soff[q], s1 = min(A.soff[q + 1], A.ns);
      const uint64_t ns = (uint64_t(s1 - s0) + A.slice - 1) / A.slice;
      const uint64_t c = local / ns, sl = local % ns;
      const uint32_t rlo = min(r1, r0 + uint32_t(c) * uint32_t(A.base));
      const uint32_t rhi = min(r1, rlo + uint32_t(A.base));
      const uint64_t slo64 = uint64_t(s0) + sl * uint64_t(A.slice);
      const uint32_t slo = uint32_t(min(uint64_t(s1), slo64));
      const uint32_t shi = uint32_t(min(uint64_t(s1), slo64 + uint64_t(A.slice)));
      s_rlo = rlo; s_rhi = rhi; s_slo = slo; s_shi = shi; s_q = uint32_t(q);
      s_fd = fastdiv_make(int32_t(2 * (rhi - rlo) + 2));
    }
    __syncthreads();
    const uint32_t rlo = s_rlo, rhi = s_rhi, slo = s_slo, shi = s_shi;
    const FastDiv fd = s_fd;
    const uint32_t tsize = uint32_t(fd.d);
    const key_t sentinel = key_t(s_q ^ 1u);
    for (uint32_t i = tid; i < tsize; i += BLOCK) s_tkeys[i] = sentinel;
    __syncthreads();
    // Build.
    for (uint32_t i = rlo + tid; i < rhi; i += BLOCK) {
      const tup_t t = rt[i];
      const key_t k = KTr::key(t);
      uint32_t slot = hash_slot(murmur3(k), fd);
      while (true) {
        const key_t old = lds_cas(&s_tkeys[slot], sentinel, k);
        if (old == sentinel) {
          s_tpays[slot] = KTr::pay(t);
          break;
        }
        slot = (slot + 1 == tsize) ? 0u : slot + 1;
      }
    }
    __syncthreads();
    // Probe, one tuple per thread per round.
    for (uint32_t r0 = slo; r0 < shi; r0 += BLOCK) {
      const uint32_t j = r0 + tid;
      bool pending = j < shi;
      key_t k = 0;
      int32_t ppay = 0;
      uint32_t slot = 0;
      if (pending) {
        const tup_t t = st[j];
        k = KTr::key(t);
        ppay = KTr::pay(t);
        slot = hash_slot(murmur3(k), fd);
      }
      while (true) {
        bool has = false;
        int32_t bpay = 0;
        if (pending) {
          while (true) {
            const key_t kk = s_tkeys[slot];
            if (kk == sentinel) { pending = false; break; }
            if (kk == k) { has = true; bpay = s_tpays[slot]; break; }
            slot = (slot + 1 == tsize) ? 0u : slot + 1;
          }
        }
        const uint64_t hm = __ballot(has);
        uint32_t wbase = 0;
        if (lane == 0 && hm) wbase = atomicAdd(&ctl.n, uint32_t(__popcll(hm)));
        wbase = __shfl(wbase, 0, kWave);
        if (has) {
          s_stage[wbase + mbcnt(hm)] = make_uint2(uint32_t(bpay), uint32_t(ppay));
          if (UNIQUE) pending = false;
          else slot = (slot + 1 == tsize) ? 0u : slot + 1;
        }
        const int any = __syncthreads_or(pending);
        const uint32_t staged = ctl.n;
        __syncthreads();  // every wave has read ctl.n before the next round adds to it
        if (staged > uint32_t(kJoinStage - BLOCK)) flush_stage<BLOCK>(A, &ctl, s_stage);
        if (!any) break;
      }
    }
    __syncthreads();
  }
  if (ctl.n > 0) flush_stage<BLOCK>(A, &ctl, s_stage);
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
  __shared__ uint32_t s_w;
  __shared__ uint64_t s_fill_sum;
  const int tid = threadIdx.x;
  const uint64_t CH = uint64_t(kOutChunk);
  if (tid == 0) { s_w = 0; s_fill_sum = 0; }
  __syncthreads();
  uint64_t rec = tid < nwg ? wg_last[tid] : ~0ull;
  s_rec[tid] = rec;
  if (rec != ~0ull) {
    atomicAdd(&s_w, 1u);
    atomicAdd(reinterpret_cast<unsigned long long*>(&s_fill_sum), (unsigned long long)(rec & 0xffffffffull));
  }
  __syncthreads();
  // rank by chunk id (ids are unique); records sort by (chunk << 32 | fill)
  if (rec != ~0ull) {
    uint32_t rank = 0;
    for (int w = 0; w < nwg; ++w) rank += (s_rec[w] != ~0ull && s_rec[w] < rec) ? 1u : 0u;
    s_sorted[rank] = rec;
  }
  __syncthreads();
  if (tid != 0) return;
  const uint64_t C = *chunk_ctr;
  const uint32_t W = s_w;
  const uint64_t total = (C - W) * CH + s_fill_sum;
  *d_count = int64_t(total);
  st->total = total;
  if (total > uint64_t(capacity)) {
    st->moves = 0;
    st->n_holes = 0;
    st->n_strays = 0;
    hole_pre[0] = 0;
    stray_pre[0] = 0;
    return;
  }
  uint32_t nh = 0, ns = 0;
  uint64_t hsum = 0, ssum = 0;
  for (uint32_t i = 0; i < W; ++i) {
    const uint64_t c = s_sorted[i] >> 32, f = s_sorted[i] & 0xffffffffull;
    const uint64_t lo = c * CH + f, hi = min((c + 1) * CH, total);
    if (lo < hi) {
      hole_lo[nh] = lo;
      hole_pre[nh] = hsum;
      hsum += hi - lo;
      ++nh;
    }
  }
  hole_pre[nh] = hsum;
  uint64_t cursor = total;
  const uint64_t end = C * CH;
  for (uint32_t i = 0; i < W; ++i) {
    const uint64_t c = s_sorted[i] >> 32, f = s_sorted[i] & 0xffffffffull;
    const uint64_t plo = c * CH + f, phi = (c + 1) * CH;
    if (phi <= total) continue;
    const uint64_t l = plo > total ? plo : total;
    if (l > cursor) {
      stray_lo[ns] = cursor;
      stray_pre[ns] = ssum;
      ssum += l - cursor;
      ++ns;
    }
    if (phi > cursor) cursor = phi;
  }
  if (end > cursor) {
    stray_lo[ns] = cursor;
    stray_pre[ns] = ssum;
    ssum += end - cursor;
    ++ns;
  }
  stray_pre[ns] = ssum;
  st->n_holes = nh;
  st->n_strays = ns;
  st->moves = hsum < ssum ? hsum : ssum;
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

__global__ __launch_bounds__(256) void k_defrag_move(const DefragState* st, const uint64_t* hole_lo,
                                                     const uint64_t* hole_pre, const uint64_t* stray_lo,
                                                     const uint64_t* stray_pre, uint2* out, int64_t capacity,
                                                     const uint2* spill) {
  const uint64_t moves = st->moves;
  const uint32_t nh = st->n_holes, ns = st->n_strays;
  for (uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; k < moves;
       k += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t si = upper_index(stray_pre, ns, k);
    const uint32_t hi = upper_index(hole_pre, nh, k);
    const uint64_t src = stray_lo[si] + (k - stray_pre[si]);
    const uint64_t dst = hole_lo[hi] + (k - hole_pre[hi]);
    const uint2 v = src < uint64_t(capacity) ? out[src] : spill[src - uint64_t(capacity)];
    out[dst] = v;
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_join_plan(const Layout& L, int base, int64_t slice, hipStream_t s) {
  hipLaunchKernelGGL(k_join_plan, dim3(1), dim3(1024), 0, s, L.t[0].off, L.t[1].off, L.pp.P, base, slice,
                     L.item_scan, L.chunk_ctr);
  return hipGetLastError();
}

static hipError_t launch_join_kernel(const JoinArgs& A, int key_bytes, int base, int unique, hipStream_t s) {
  const int slots = 2 * base + 2;  // table of a chunk of c tuples has 2c + 2 slots
  JoinArgs a = A;
  a.table_cap = slots;
  const size_t lds = join_lds_bytes(key_bytes, slots);
  dim3 grid(kJoinGrid), block(kJoinBlock);
  if (key_bytes == 4) {
    if (unique) hipLaunchKernelGGL((k_join<4, kJoinBlock, true>), grid, block, lds, s, a);
    else hipLaunchKernelGGL((k_join<4, kJoinBlock, false>), grid, block, lds, s, a);
  } else {
    if (unique) hipLaunchKernelGGL((k_join<8, kJoinBlock, true>), grid, block, lds, s, a);
    else hipLaunchKernelGGL((k_join<8, kJoinBlock, false>), grid, block, lds, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_join(const Layout& L, int key_bytes, int base, int64_t slice, int unique, pj_pair* out,
                       int64_t capacity, hipStream_t s) {
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
  return launch_join_kernel(A, key_bytes, base, unique, s);
}

hipError_t launch_defrag(const Layout& L, pj_pair* out, int64_t capacity, int64_t* d_count, hipStream_t s) {
  hipLaunchKernelGGL(k_defrag_plan, dim3(1), dim3(1024), 0, s, L.chunk_ctr, L.wg_last, kJoinGrid, capacity, d_count,
                     L.dstate, L.hole_lo, L.hole_pre, L.stray_lo, L.stray_pre);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_defrag_move, dim3(512), dim3(256), 0, s, L.dstate, L.hole_lo, L.hole_pre, L.stray_lo,
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
  return b + 1024;
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
  hipError_t e = launch_join_plan(L, base, slice, s);
  if (e != hipSuccess) return e;
  e = launch_join(L, key_bytes, base, slice, unique, out, capacity, s);
  if (e != hipSuccess) return e;
  return launch_defrag(L, out, capacity, d_count, s);
}

}  // namespace pj
