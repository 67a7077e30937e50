// pj_api.cpp -- host orchestration and the extern "C" boundary of libpj.so.
//
// The reference drives its phases from main.cu:759-893 with blocking copies and
// per-phase event synchronisation; here one stream-ordered sequence of kernels
// runs with no host round trip, on caller-provided workspace.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/pj.h"
#include "pj_device.h"
#include "pj_internal.h"
#include "pj_misc.h"
#include "pj_planner.h"

namespace pj {

static thread_local std::string g_err;

static pj_status fail(pj_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return st;
}

#define PJ_HIP(call)                                                                             \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) return fail(PJ_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

int tuple_bytes(int key_bytes) { return key_bytes == 4 ? 8 : 16; }

static int scatter_tile(int key_bytes) { return key_bytes == 4 ? kScatterTile4 : kScatterTile8; }

uint32_t choose_nseg(int64_t n) {
  int64_t s = n / (int64_t(kNumXcd) * 2 * kScatterTile4);
  if (s < 1) s = 1;
  if (s > kMaxSegPerXcd) s = kMaxSegPerXcd;
  return uint32_t(s) * kNumXcd;
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct Carver {
  char* base;
  size_t off = 0;
  void* take(size_t bytes) {
    void* p = base ? base + off : nullptr;
    off += align256(bytes);
    return p;
  }
};

bool deep_plan(const PartParams& pp) { return pp.passes == 2 && pp.B > kMaxCombinedBits; }

// Deep plans keep only the [segment][pass-1 digit] matrix (in mat) and count
// the pass-2 digits per pass-2 item (h2) after pass 1.
static void carve_table(Carver& c, int64_t n, int key_bytes, const PartParams& pp, bool two_pass, PartTable* T) {
  const size_t tb = size_t(tuple_bytes(key_bytes));
  const bool deep = deep_plan(pp);
  T->n = uint32_t(n);
  T->nseg = choose_nseg(n);
  T->buf1 = c.take(size_t(n) * tb + 64);
  T->buf2 = two_pass ? c.take(size_t(n) * tb + 64) : nullptr;
  T->mat = static_cast<uint32_t*>(c.take(size_t(T->nseg) * (deep ? pp.F1 : pp.P) * 4));
  T->matc = two_pass && !deep ? static_cast<uint32_t*>(c.take(size_t(T->nseg) * pp.F1 * 4)) : nullptr;
  T->tot = static_cast<uint32_t*>(c.take(size_t(pp.P) * 4));
  T->totc = two_pass ? static_cast<uint32_t*>(c.take(size_t(pp.F1) * 4)) : nullptr;
  T->off = static_cast<uint32_t*>(c.take(size_t(pp.P + 1) * 4));
  T->offc = T->h2 = T->ipre = nullptr;
  if (deep) {
    T->offc = static_cast<uint32_t*>(c.take(size_t(pp.F1 + 1) * 4));
    T->h2 = static_cast<uint32_t*>(c.take(size_t(deep_item_capacity(n, key_bytes, pp)) * pp.F2 * 4));
    T->ipre = static_cast<uint32_t*>(c.take(size_t((pp.F1 + kNumXcd - 1) / kNumXcd * kNumXcd + 1) * 4));
  }
}

static void carve_join(Carver& c, const PartParams& pp, uint64_t desc_cap, Layout* L) {
  L->item_scan = static_cast<uint64_t*>(c.take(size_t(pp.P + 1) * 8));
  L->chunk_ctr = static_cast<uint32_t*>(c.take(64));
  L->dstate = static_cast<DefragState*>(c.take(sizeof(DefragState)));
  L->wg_last = static_cast<uint64_t*>(c.take(size_t(kJoinGrid) * 8));
  L->hole_lo = static_cast<uint64_t*>(c.take((size_t(kJoinGrid) + 3) * 8));
  L->hole_pre = static_cast<uint64_t*>(c.take((size_t(kJoinGrid) + 3) * 8));
  L->stray_lo = static_cast<uint64_t*>(c.take((size_t(kJoinGrid) + 3) * 8));
  L->stray_pre = static_cast<uint64_t*>(c.take((size_t(kJoinGrid) + 3) * 8));
  L->spill = static_cast<pj_pair*>(c.take(size_t(kJoinGrid) * size_t(kOutChunk) * 8));
  L->scan_scratch = c.take(2 * 4096 * 8);
  L->desc_cap = desc_cap;
  L->desc = c.take(size_t(desc_cap) * 32);
}

// Work item descriptors: enough for every partition plus every extra build
// chunk and probe slice of a skewed table; items past the bound (only when
// heavy keys sit on both sides) are decoded by search instead.
static uint64_t desc_capacity(int64_t nR, int64_t nS, int P, int base, int64_t slice) {
  return uint64_t(P) + uint64_t(nR / base) + uint64_t(nS / slice) + 4096;
}

size_t layout_join(int64_t nR, int64_t nS, int key_bytes, const PartParams& pp, int jbase, int64_t slice, char* base,
                   Layout* L) {
  Carver c{base};
  Layout tmp{};
  Layout* Lp = L ? L : &tmp;
  Lp->ntables = 2;
  Lp->pp = pp;
  carve_table(c, nR, key_bytes, pp, pp.passes == 2, &Lp->t[0]);
  carve_table(c, nS, key_bytes, pp, pp.passes == 2, &Lp->t[1]);
  carve_join(c, pp, desc_capacity(nR, nS, pp.P, jbase, slice), Lp);
  c.take(64);  // device pair count of the synchronous entry point
  Lp->bytes = c.off;
  return c.off;
}

// ---------------------------------------------------------------------------
// Validation and plans
// ---------------------------------------------------------------------------
static pj_status check_table(const pj_table* t, const char* name, bool need_data = true) {
  if (!t) return fail(PJ_ERR_INVALID_ARG, "%s table is NULL", name);
  if (t->key_bytes != 4 && t->key_bytes != 8)
    return fail(PJ_ERR_INVALID_ARG, "%s key_bytes must be 4 or 8 (got %d)", name, t->key_bytes);
  if (t->n < 0 || t->n >= (int64_t(1) << 31))
    return fail(PJ_ERR_INVALID_ARG, "%s row count %lld outside [0, 2^31)", name, (long long)t->n);
  if (need_data && t->n > 0 && !t->keys) return fail(PJ_ERR_INVALID_ARG, "%s keys pointer is NULL", name);
  return PJ_OK;
}

// need_data = false for the planning functions, which read only sizes and widths.
static pj_status check_pair(const pj_table* b, const pj_table* p, bool need_data = true) {
  pj_status st = check_table(b, "build", need_data);
  if (st != PJ_OK) return st;
  st = check_table(p, "probe", need_data);
  if (st != PJ_OK) return st;
  if (b->key_bytes != p->key_bytes)
    return fail(PJ_ERR_INVALID_ARG, "build and probe key widths differ (%d vs %d)", b->key_bytes, p->key_bytes);
  return PJ_OK;
}

// Build chunks of f = 1.5 times the average partition: the reference defaults to
// 3.0 (main.cu:96) to absorb skew inside one block, here heavy partitions are
// split on both sides instead, and the smaller table lets two join workgroups
// share a CU's LDS.
// Build tuples per partition, on average, that the plain plan aims at: 2048
// for 4-byte keys, 1024 for 8-byte keys (build chunks hold at most
// join_table_capacity / 2 - 1 tuples: 2783 and 2047).
static double partition_target(int key_bytes) { return key_bytes == 4 ? 2048.0 : 1024.0; }

static void plain_plan(int64_t nR, int key_bytes, pj_plan* plan) {
  std::memset(plan, 0, sizeof(*plan));
  // Build chunks of up to 1.25x the average partition: 2560 tuples at the
  // 4-byte-key target, five per join thread (at 1.5x the chunk needs six, and
  // that kernel's registers spill without the unique-key promise: 64M x 256M
  // 3.81 against 3.68 ms).
  const double f = 1.25;
  // The fewest radix bits whose partitions hold at most partition_target()
  // build tuples on average.  Fewer, larger
  // partitions make the histogram, the offsets and the scatters cheaper than
  // they make the join dearer while the chunks still fit: 64M x 256M int32
  // 3.61 ms in 8 + 8 bits, 3.47 ms in 8 + 7; 16M x 256M Zipf(0.5) 3.21 ms in
  // 7 + 7, 3.04 ms in 7 + 6; with 8-byte keys 2048-tuple partitions overflow
  // the chunk (64M x 256M: 7.17 ms in 8 + 8, 8.43 ms in 8 + 7).
  const double T = partition_target(key_bytes);
  int B = 1;
  while (B < kMaxCombinedBits && double(nR) / double(int64_t(1) << B) > T) ++B;
  plan->max_partition_size_factor = f;
  plan->unique_build_keys = 0;
  plan->probe_slice = 0;
  plan->source = 0;
  // One pass up to 1024 partitions; from 2048 on, two passes beat one scatter
  // over that many runs (int64 2M x 8M: 0.50 ms in one pass of 11 bits, 0.42 ms
  // in 6 + 5; at 10 bits one pass stays ahead, 0.22 against 0.27 ms for 1M x 4M).
  if (B <= 10) {
    plan->passes = 1;
    plan->bits1 = B;
    plan->bits2 = 0;
  } else {
    plan->passes = 2;
    plan->bits1 = (B + 1) / 2;
    plan->bits2 = B / 2;
  }
}

static pj_status check_plan(const pj_plan* plan) {
  if (plan->passes != 1 && plan->passes != 2)
    return fail(PJ_ERR_INVALID_ARG, "plan.passes must be 1 or 2 (got %d)", plan->passes);
  if (plan->bits1 < 1 || plan->bits1 > 12)
    return fail(PJ_ERR_INVALID_ARG, "plan.bits1 must be in [1, 12] (got %d)", plan->bits1);
  if (plan->passes == 2 && (plan->bits2 < 1 || plan->bits2 > 12))
    return fail(PJ_ERR_INVALID_ARG, "plan.bits2 must be in [1, 12] for two passes (got %d)", plan->bits2);
  if (plan->passes == 1 && plan->bits2 != 0)
    return fail(PJ_ERR_INVALID_ARG, "plan.bits2 must be 0 for one pass (got %d)", plan->bits2);
  if (plan->bits1 + plan->bits2 > kMaxTotalBits)
    return fail(PJ_ERR_UNSUPPORTED, "bits1 + bits2 = %d exceeds %d", plan->bits1 + plan->bits2, kMaxTotalBits);
  if (!(plan->max_partition_size_factor > 0.0))
    return fail(PJ_ERR_INVALID_ARG, "plan.max_partition_size_factor must be > 0");
  if (plan->probe_slice < 0) return fail(PJ_ERR_INVALID_ARG, "plan.probe_slice must be >= 0");
  return PJ_OK;
}

static PartParams params_of(const pj_plan* plan) {
  PartParams pp{};
  pp.passes = plan->passes;
  pp.b1 = plan->bits1;
  pp.b2 = plan->passes == 2 ? plan->bits2 : 0;
  pp.B = pp.b1 + pp.b2;
  pp.P = 1 << pp.B;
  pp.F1 = 1 << pp.b1;
  pp.F2 = 1 << pp.b2;
  pp.shift0 = 0;
  return pp;
}

// base = ceil(f * |R| / P) (main.cu:867), clamped to the LDS table.
static int join_base(int64_t nR, int P, double f, int key_bytes) {
  double b = std::ceil(f * double(nR) / double(P));
  const int cap = join_table_capacity(key_bytes) / 2 - 1;
  if (b < 1.0) b = 1.0;
  if (b > double(cap)) b = double(cap);
  return int(b);
}

static int64_t join_slice(const pj_plan* plan, int base) {
  if (plan->probe_slice > 0) return plan->probe_slice < (int64_t(1) << 30) ? plan->probe_slice : int64_t(1) << 30;
  int64_t s = int64_t(base) * 8;
  return s < 16384 ? 16384 : s;
}

// ---------------------------------------------------------------------------
// Profiling
// ---------------------------------------------------------------------------
// A ring of event sets, one per profiled join, so that joins can be queued
// back to back and their phase times read after one wait at the end.
constexpr int kProfRing = 128;
struct Prof {
  bool on = false;
  bool recorded = false;  // the last join on this thread recorded its events
  int head = 0;           // the set the next join records into
  int count = 0;          // recorded sets since profiling was enabled (at most kProfRing)
  uint32_t marks = 0xFFu; // the phase boundaries recorded (bit i: event i)
  hipEvent_t ev[kProfRing][8] = {};
  uint8_t set_marks[kProfRing] = {};  // the boundaries each set recorded
  bool created = false;
};
static thread_local Prof g_prof;

static void prof_mark(int i, hipStream_t s) {
  if (!g_prof.on || !g_prof.created || !((g_prof.marks >> i) & 1u)) return;
  (void)hipEventRecord(g_prof.ev[g_prof.head][i], s);
}

// ---------------------------------------------------------------------------
// The join pipeline
// ---------------------------------------------------------------------------
static pj_status run_join(const pj_table* build, const pj_table* probe, const pj_plan* plan_in, pj_pair* out,
                          int64_t capacity, int64_t* d_count, void* workspace, size_t ws_bytes, hipStream_t s) {
  pj_status st = check_pair(build, probe);
  if (st != PJ_OK) return st;
  if (capacity < 0) return fail(PJ_ERR_INVALID_ARG, "out_capacity must be >= 0");
  if (capacity > 0 && !out) return fail(PJ_ERR_INVALID_ARG, "out is NULL with out_capacity > 0");
  if (!d_count) return fail(PJ_ERR_INVALID_ARG, "pair count pointer is NULL");
  pj_plan plan;
  if (plan_in) plan = *plan_in;
  else plain_plan(build->n, build->key_bytes, &plan);
  st = check_plan(&plan);
  if (st != PJ_OK) return st;
  const int kb = build->key_bytes;
  if (build->n == 0 || probe->n == 0) {
    PJ_HIP(launch_fill_u64(reinterpret_cast<uint64_t*>(d_count), 0, 1, s));
    if (g_prof.on) g_prof.recorded = false;
    return PJ_OK;
  }
  const PartParams pp = params_of(&plan);
  const int base = join_base(build->n, pp.P, plan.max_partition_size_factor, kb);
  const int64_t slice = join_slice(&plan, base);
  const size_t need = layout_join(build->n, probe->n, kb, pp, base, slice, nullptr, nullptr);
  if (!workspace || ws_bytes < need)
    return fail(PJ_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu required", ws_bytes, need);
  if (reinterpret_cast<uintptr_t>(workspace) % 256 != 0)
    return fail(PJ_ERR_INVALID_ARG, "workspace must be 256-byte aligned");
  Layout L{};
  layout_join(build->n, probe->n, kb, pp, base, slice, static_cast<char*>(workspace), &L);
  L.t[0].keys = build->keys;
  L.t[0].pay = build->payload;
  L.t[1].keys = probe->keys;
  L.t[1].pay = probe->payload;

  if (g_prof.on && !g_prof.created) {
    for (auto& set : g_prof.ev)
      for (auto& e : set) PJ_HIP(hipEventCreate(&e));
    g_prof.created = true;
  }
  prof_mark(0, s);
  if (deep_plan(pp)) {
    // Pass 1 as a one-pass partition on bits [b2, b2 + b1) into offc, then
    // the pass-2 digit counts per item over its output.
    Layout L1 = L;
    L1.pp = PartParams{pp.b1, pp.b1, 0, pp.F1, pp.F1, 1, 1, pp.b2};
    for (int t = 0; t < 2; ++t) {
      L1.t[t].off = L.t[t].offc;
      L1.t[t].tot = L.t[t].totc;
    }
    PJ_HIP(launch_hist(L1, kb, s));
    prof_mark(1, s);
    PJ_HIP(launch_colscan(L1, s));
    PJ_HIP(launch_offsets(L1, s));
    prof_mark(2, s);
    PJ_HIP(launch_pass1(L1, kb, s));
    prof_mark(3, s);
    PJ_HIP(launch_hist2(L, kb, s));
    PJ_HIP(launch_segscan(L, s));
    PJ_HIP(launch_offsets(L, s));
    PJ_HIP(launch_pass2(L, kb, s));
  } else {
    PJ_HIP(launch_hist(L, kb, s));
    prof_mark(1, s);
    PJ_HIP(launch_colscan(L, s));
    PJ_HIP(launch_offsets(L, s));
    prof_mark(2, s);
    PJ_HIP(launch_pass1(L, kb, s));
    prof_mark(3, s);
    if (pp.passes == 2) PJ_HIP(launch_pass2(L, kb, s));
  }
  prof_mark(4, s);
  PJ_HIP(launch_join_plan(L, base, slice, s));
  prof_mark(5, s);
  PJ_HIP(launch_join(L, kb, base, slice, plan.unique_build_keys, out, capacity, s));
  prof_mark(6, s);
  PJ_HIP(launch_defrag(L, out, capacity, d_count, s));
  prof_mark(7, s);
  if (g_prof.on) {
    g_prof.recorded = true;
    g_prof.set_marks[g_prof.head] = uint8_t(g_prof.marks);
    g_prof.head = (g_prof.head + 1) % kProfRing;
    g_prof.count = g_prof.count < kProfRing ? g_prof.count + 1 : kProfRing;
  }
  return PJ_OK;
}

}  // namespace pj

using namespace pj;

extern "C" {

const char* pj_version(void) { return "partitioned_join_amd 0.1 (gfx950)"; }

const char* pj_last_error(void) { return g_err.c_str(); }

pj_status pj_plan_default(const pj_table* build, const pj_table* probe, pj_plan* plan) {
  pj_status st = check_pair(build, probe, false);
  if (st != PJ_OK) return st;
  if (!plan) return fail(PJ_ERR_INVALID_ARG, "plan is NULL");
  plain_plan(build->n, build->key_bytes, plan);
  return PJ_OK;
}

pj_status pj_workspace_size(const pj_table* build, const pj_table* probe, const pj_plan* plan_in, size_t* bytes) {
  pj_status st = check_pair(build, probe, false);
  if (st != PJ_OK) return st;
  if (!bytes) return fail(PJ_ERR_INVALID_ARG, "bytes is NULL");
  pj_plan plan;
  if (plan_in) plan = *plan_in;
  else plain_plan(build->n, build->key_bytes, &plan);
  st = check_plan(&plan);
  if (st != PJ_OK) return st;
  const PartParams pp = params_of(&plan);
  const int base = join_base(build->n, pp.P, plan.max_partition_size_factor, build->key_bytes);
  *bytes = layout_join(build->n, probe->n, build->key_bytes, pp, base, join_slice(&plan, base), nullptr, nullptr);
  return PJ_OK;
}

pj_status pj_partitioned_join_async(const pj_table* build, const pj_table* probe, const pj_plan* plan, pj_pair* out,
                                    int64_t out_capacity, int64_t* d_out_count, void* workspace,
                                    size_t workspace_bytes, hipStream_t stream) {
  return run_join(build, probe, plan, out, out_capacity, d_out_count, workspace, workspace_bytes, stream);
}

pj_status pj_partitioned_join(const pj_table* build, const pj_table* probe, const pj_plan* plan, pj_pair* out,
                              int64_t out_capacity, int64_t* out_count, void* workspace, size_t workspace_bytes,
                              hipStream_t stream) {
  if (!out_count) return fail(PJ_ERR_INVALID_ARG, "out_count is NULL");
  pj_status st = check_pair(build, probe);
  if (st != PJ_OK) return st;
  // The device pair count lives in the last 64 bytes of the layout.
  size_t need = 0;
  st = pj_workspace_size(build, probe, plan, &need);
  if (st != PJ_OK) return st;
  if (!workspace || workspace_bytes < need)
    return fail(PJ_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu required", workspace_bytes, need);
  int64_t* d_count = reinterpret_cast<int64_t*>(static_cast<char*>(workspace) + need - 256);
  st = run_join(build, probe, plan, out, out_capacity, d_count, workspace, workspace_bytes, stream);
  if (st != PJ_OK) return st;
  int64_t h = 0;
  PJ_HIP(hipMemcpyAsync(&h, d_count, sizeof(h), hipMemcpyDeviceToHost, stream));
  PJ_HIP(hipStreamSynchronize(stream));
  *out_count = h;
  if (h > out_capacity)
    return fail(PJ_ERR_CAPACITY, "join produced %lld pairs, output capacity is %lld", (long long)h,
                (long long)out_capacity);
  return PJ_OK;
}

// ---------------------------------------------------------------------------
// The allocation-free join: the workspace is a file-scope device array, one
// copy per GPU, held by the loaded code object until the process ends.
// ---------------------------------------------------------------------------
}  // extern "C"

namespace pj {

constexpr size_t kNoAllocScratchBytes = size_t(192) << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// The device address of the current device's copy of `scratch` (the array's
// name in host code is only a host-side stand-in), looked up once per device.
static pj_status scratch_address(void** ptr) {
  int dev = -1;
  PJ_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(PJ_ERR_UNSUPPORTED, "device index %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    PJ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p || reinterpret_cast<uintptr_t>(p) % 256 != 0)
      return fail(PJ_ERR_HIP, "hipGetSymbolAddress returned an unusable scratch address");
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return PJ_OK;
}

}  // namespace pj

extern "C" {

pj_status pj_partitioned_join_noalloc(const pj_table* build, const pj_table* probe, const pj_plan* plan,
                                      pj_pair* out, int64_t out_capacity, int64_t* d_out_count, hipStream_t stream) {
  size_t need = 0;
  pj_status st = pj_workspace_size(build, probe, plan, &need);
  if (st != PJ_OK) return st;
  if (need > kNoAllocScratchBytes)
    return fail(PJ_ERR_WORKSPACE,
                "this join needs %zu bytes of workspace; the allocation-free entry point has %zu (use "
                "pj_partitioned_join_async with a caller workspace)",
                need, kNoAllocScratchBytes);
  void* ws = nullptr;
  st = scratch_address(&ws);
  if (st != PJ_OK) return st;
  return run_join(build, probe, plan, out, out_capacity, d_out_count, ws, kNoAllocScratchBytes, stream);
}

pj_status pj_project_reference(const pj_table* build, const pj_pair* pairs, int64_t n, void* out_xy,
                               hipStream_t stream) {
  pj_status st = check_table(build, "build");
  if (st != PJ_OK) return st;
  if (n < 0 || (n > 0 && (!pairs || !out_xy))) return fail(PJ_ERR_INVALID_ARG, "bad pairs/out_xy/n");
  if (n == 0) return PJ_OK;
  PJ_HIP(project_reference(build->keys, build->key_bytes, pairs, n, out_xy, stream));
  return PJ_OK;
}

pj_status pj_gather(const pj_pair* pairs, int64_t n, const void* build_col, int32_t build_bytes,
                    const void* probe_col, int32_t probe_bytes, void* out_build, void* out_probe,
                    hipStream_t stream) {
  if (n < 0 || (n > 0 && !pairs)) return fail(PJ_ERR_INVALID_ARG, "bad pairs/n");
  if (build_col && build_bytes != 4 && build_bytes != 8)
    return fail(PJ_ERR_INVALID_ARG, "build column width must be 4 or 8 (got %d)", build_bytes);
  if (probe_col && probe_bytes != 4 && probe_bytes != 8)
    return fail(PJ_ERR_INVALID_ARG, "probe column width must be 4 or 8 (got %d)", probe_bytes);
  if (n == 0 || (!build_col && !probe_col)) return PJ_OK;
  if ((build_col && !out_build) || (probe_col && !out_probe))
    return fail(PJ_ERR_INVALID_ARG, "a gathered column needs an output buffer");
  PJ_HIP(gather_columns(pairs, n, build_col, build_col ? build_bytes : 4, probe_col, probe_col ? probe_bytes : 4,
                        out_build, out_probe, stream));
  return PJ_OK;
}

// ---------------------------------------------------------------------------
// Phase-level entry points
// ---------------------------------------------------------------------------
static void one_table_layout(Carver& c, const pj_table* t, const PartParams& pp, Layout* L) {
  std::memset(L, 0, sizeof(*L));
  L->ntables = 1;
  L->pp = pp;
  PartTable& T = L->t[0];
  T.keys = t->keys;
  T.pay = t->payload;
  T.n = uint32_t(t->n);
  T.nseg = choose_nseg(t->n);
  T.mat = static_cast<uint32_t*>(c.take(size_t(T.nseg) * pp.P * 4));
  T.matc = pp.passes == 2 ? static_cast<uint32_t*>(c.take(size_t(T.nseg) * pp.F1 * 4)) : nullptr;
  L->scan_scratch = c.take(2 * 4096 * 8);
  L->t[1] = T;  // unused second slot mirrors the first
}

size_t pj_histogram_workspace_size(int32_t total_bits) {
  if (total_bits < 1 || total_bits > kMaxCombinedBits) return 0;
  const size_t P = size_t(1) << total_bits;
  return size_t(kMaxSegPerXcd * kNumXcd) * (P + 4096) * 4 + 4 * 256 + P * 4 + 2 * 4096 * 8;
}

pj_status pj_combined_histogram(const pj_table* table, int32_t total_bits, int32_t bits2, uint32_t* d_fine,
                                uint32_t* d_pass1, void* workspace, size_t workspace_bytes, hipStream_t stream) {
  pj_status st = check_table(table, "input");
  if (st != PJ_OK) return st;
  if (total_bits < 1 || total_bits > kMaxCombinedBits) return fail(PJ_ERR_INVALID_ARG, "total_bits must be in [1, 16]");
  if (bits2 < 0 || total_bits - bits2 < 1 || total_bits - bits2 > 12)
    return fail(PJ_ERR_INVALID_ARG, "pass-1 bits (total_bits - bits2) must be in [1, 12]");
  if (!d_fine || !d_pass1) return fail(PJ_ERR_INVALID_ARG, "output histogram pointer is NULL");
  if (workspace_bytes < pj_histogram_workspace_size(total_bits) || !workspace)
    return fail(PJ_ERR_WORKSPACE, "histogram workspace too small");
  PartParams pp{};
  pp.B = total_bits;
  pp.b2 = bits2;
  pp.b1 = total_bits - bits2;
  pp.P = 1 << total_bits;
  pp.F1 = 1 << pp.b1;
  pp.F2 = 1 << bits2;
  pp.passes = bits2 > 0 ? 2 : 1;
  Carver c{static_cast<char*>(workspace)};
  Layout L;
  one_table_layout(c, table, pp, &L);
  L.t[0].tot = d_fine;
  L.t[0].totc = d_pass1;
  L.t[1] = L.t[0];
  if (table->n == 0) {
    PJ_HIP(launch_fill_u32(d_fine, 0, pp.P, stream));
    PJ_HIP(launch_fill_u32(d_pass1, 0, pp.F1, stream));
    return PJ_OK;
  }
  PJ_HIP(launch_hist(L, table->key_bytes, stream));
  PJ_HIP(launch_colscan(L, stream));
  if (pp.passes == 1) PJ_HIP(hipMemcpyAsync(d_pass1, d_fine, size_t(pp.P) * 4, hipMemcpyDeviceToDevice, stream));
  return PJ_OK;
}

pj_status pj_histogram(const pj_table* table, int32_t shift, int32_t bits, uint32_t* d_hist, void* workspace,
                       size_t workspace_bytes, hipStream_t stream) {
  (void)workspace;
  (void)workspace_bytes;
  pj_status st = check_table(table, "input");
  if (st != PJ_OK) return st;
  if (bits < 1 || bits > 15) return fail(PJ_ERR_INVALID_ARG, "bits must be in [1, 15]");
  if (shift < 0 || shift + bits > 8 * table->key_bytes) return fail(PJ_ERR_INVALID_ARG, "bad shift");
  if (shift + bits > 32) return fail(PJ_ERR_UNSUPPORTED, "digits above bit 31 are not supported");
  if (!d_hist) return fail(PJ_ERR_INVALID_ARG, "d_hist is NULL");
  PJ_HIP(digit_hist(table->keys, table->n, table->key_bytes, 1, shift, bits, d_hist, stream));
  return PJ_OK;
}

size_t pj_scan_workspace_size(int64_t n) { return scan_ws_bytes(n); }

static pj_status scan_common(const uint32_t* in, uint32_t* out, int64_t n, void* ws, size_t wsb, hipStream_t s,
                             int inclusive) {
  if (n < 0 || (n > 0 && (!in || !out))) return fail(PJ_ERR_INVALID_ARG, "bad scan arguments");
  if (n > (int64_t(1) << 34)) return fail(PJ_ERR_UNSUPPORTED, "scan length too large");
  if (n == 0) return PJ_OK;
  if (!ws || wsb < scan_ws_bytes(n)) return fail(PJ_ERR_WORKSPACE, "scan workspace too small");
  PJ_HIP(scan_u32(in, out, n, inclusive, static_cast<uint32_t*>(ws), s));
  return PJ_OK;
}

pj_status pj_exclusive_scan(const uint32_t* d_in, uint32_t* d_out, int64_t n, void* workspace, size_t workspace_bytes,
                            hipStream_t stream) {
  return scan_common(d_in, d_out, n, workspace, workspace_bytes, stream, 0);
}

pj_status pj_inclusive_scan(const uint32_t* d_in, uint32_t* d_out, int64_t n, void* workspace, size_t workspace_bytes,
                            hipStream_t stream) {
  return scan_common(d_in, d_out, n, workspace, workspace_bytes, stream, 1);
}

size_t pj_partition_workspace_size(const pj_table* table, int32_t bits) {
  if (!table || bits < 1 || bits > 12 || table->n < 0) return 0;
  const size_t P = size_t(1) << bits;
  return size_t(choose_nseg(table->n)) * P * 4 + P * 4 + 2 * 4096 * 8 + 1024;
}

pj_status pj_partition_pass(const pj_table* table, int32_t shift, int32_t bits, void* d_tuples_out,
                            uint32_t* d_offsets, void* workspace, size_t workspace_bytes, hipStream_t stream) {
  pj_status st = check_table(table, "input");
  if (st != PJ_OK) return st;
  if (bits < 1 || bits > 12) return fail(PJ_ERR_INVALID_ARG, "bits must be in [1, 12]");
  if (shift < 0 || shift + bits > 32) return fail(PJ_ERR_INVALID_ARG, "shift + bits must be <= 32");
  if (!d_offsets || (table->n > 0 && !d_tuples_out)) return fail(PJ_ERR_INVALID_ARG, "NULL output");
  if (!workspace || workspace_bytes < pj_partition_workspace_size(table, bits))
    return fail(PJ_ERR_WORKSPACE, "partition workspace too small");
  PartParams pp{};
  pp.B = bits;
  pp.b1 = bits;
  pp.b2 = 0;
  pp.P = 1 << bits;
  pp.F1 = pp.P;
  pp.F2 = 1;
  pp.passes = 1;
  pp.shift0 = shift;
  Carver c{static_cast<char*>(workspace)};
  Layout L;
  one_table_layout(c, table, pp, &L);
  L.t[0].tot = static_cast<uint32_t*>(c.take(size_t(pp.P) * 4));
  L.t[0].off = d_offsets;
  L.t[0].buf1 = d_tuples_out;
  L.t[1] = L.t[0];
  PJ_HIP(launch_hist(L, table->key_bytes, stream));
  PJ_HIP(launch_colscan(L, stream));
  PJ_HIP(launch_offsets(L, stream));
  if (table->n > 0) PJ_HIP(launch_pass1(L, table->key_bytes, stream));
  return PJ_OK;
}

size_t pj_partition2_workspace_size(int64_t n, int32_t bits1, int32_t bits2) {
  (void)n;
  if (bits1 < 1 || bits1 > 12 || bits2 < 1 || bits2 > 12 || bits1 + bits2 > kMaxCombinedBits) return 0;
  const size_t P = size_t(1) << (bits1 + bits2), F1 = size_t(1) << bits1;
  return size_t(kNumXcd) * (P + F1) * 4 + P * 4 + 2 * 4096 * 8 + 4 * 256;
}

pj_status pj_partition_pass2(const void* d_tuples_in, int64_t n, int32_t key_bytes, int32_t bits1, int32_t bits2,
                             const uint32_t* d_offsets1, void* d_tuples_out, uint32_t* d_offsets2, void* workspace,
                             size_t workspace_bytes, hipStream_t stream) {
  if (key_bytes != 4 && key_bytes != 8) return fail(PJ_ERR_INVALID_ARG, "key_bytes must be 4 or 8");
  if (n < 0 || n >= (int64_t(1) << 31)) return fail(PJ_ERR_INVALID_ARG, "n outside [0, 2^31)");
  if (bits1 < 1 || bits1 > 12 || bits2 < 1 || bits2 > 12 || bits1 + bits2 > kMaxCombinedBits)
    return fail(PJ_ERR_INVALID_ARG, "bits1, bits2 must be in [1, 12] with bits1 + bits2 <= 16");
  if (!d_offsets1 || !d_offsets2 || (n > 0 && (!d_tuples_in || !d_tuples_out)))
    return fail(PJ_ERR_INVALID_ARG, "NULL argument");
  if (!workspace || workspace_bytes < pj_partition2_workspace_size(n, bits1, bits2))
    return fail(PJ_ERR_WORKSPACE, "pass-2 workspace too small");
  PartParams pp{};
  pp.b1 = bits1;
  pp.b2 = bits2;
  pp.B = bits1 + bits2;
  pp.P = 1 << pp.B;
  pp.F1 = 1 << bits1;
  pp.F2 = 1 << bits2;
  pp.passes = 2;
  Carver c{static_cast<char*>(workspace)};
  Layout L{};
  L.ntables = 1;
  L.pp = pp;
  PartTable& T = L.t[0];
  T.n = uint32_t(n);
  T.nseg = kNumXcd;
  T.buf1 = const_cast<void*>(d_tuples_in);
  T.buf2 = d_tuples_out;
  T.mat = static_cast<uint32_t*>(c.take(size_t(kNumXcd) * pp.P * 4));
  T.matc = static_cast<uint32_t*>(c.take(size_t(kNumXcd) * pp.F1 * 4));
  T.tot = static_cast<uint32_t*>(c.take(size_t(pp.P) * 4));
  T.off = d_offsets2;
  L.scan_scratch = c.take(2 * 4096 * 8);
  L.t[1] = T;
  PJ_HIP(digit_hist(d_tuples_in, n, key_bytes, 2, 0, pp.B, T.tot, stream));
  PJ_HIP(launch_offsets(L, stream));
  PJ_HIP(p2_prep(d_offsets1, pp.F1, T.nseg, T.matc, T.mat, pp.P, stream));
  if (n > 0) PJ_HIP(launch_pass2(L, key_bytes, stream));
  return PJ_OK;
}

size_t pj_assign_blocks_workspace_size(int32_t num_partitions) {
  return num_partitions < 1 ? 0 : scan_ws_bytes(num_partitions);
}

pj_status pj_assign_blocks(const uint32_t* d_offsets, int64_t total, int32_t num_partitions, int32_t base,
                           uint32_t* d_blocks_scanned, uint32_t* d_partition_index, int64_t max_blocks,
                           uint32_t* d_total_blocks, void* workspace, size_t workspace_bytes, hipStream_t stream) {
  if (num_partitions < 1 || base < 1 || total < 0 || total >= (int64_t(1) << 31) || max_blocks < 0)
    return fail(PJ_ERR_INVALID_ARG, "bad assign_blocks sizes");
  if (!d_offsets || !d_blocks_scanned || !d_total_blocks || (max_blocks > 0 && !d_partition_index))
    return fail(PJ_ERR_INVALID_ARG, "NULL argument");
  if (!workspace || workspace_bytes < pj_assign_blocks_workspace_size(num_partitions))
    return fail(PJ_ERR_WORKSPACE, "assign_blocks workspace too small");
  PJ_HIP(assign_blocks(d_offsets, total, num_partitions, base, d_blocks_scanned, d_partition_index, max_blocks,
                       d_total_blocks, static_cast<uint32_t*>(workspace), stream));
  return PJ_OK;
}

size_t pj_join_workspace_size(int32_t total_bits) {
  if (total_bits < 1 || total_bits > kMaxCombinedBits) return 0;
  return join_generic_ws_bytes(total_bits) + 4096;
}

pj_status pj_join_partitions(const void* d_build_tuples, const uint32_t* d_build_offsets, const void* d_probe_tuples,
                             const uint32_t* d_probe_offsets, int32_t key_bytes, int32_t total_bits, int32_t base,
                             int64_t probe_slice, int32_t unique_build_keys, pj_pair* out, int64_t out_capacity,
                             int64_t* d_out_count, void* workspace, size_t workspace_bytes, hipStream_t stream) {
  if (key_bytes != 4 && key_bytes != 8) return fail(PJ_ERR_INVALID_ARG, "key_bytes must be 4 or 8");
  if (total_bits < 1 || total_bits > kMaxCombinedBits) return fail(PJ_ERR_INVALID_ARG, "total_bits must be in [1, 16]");
  if (!d_build_offsets || !d_probe_offsets || !d_out_count) return fail(PJ_ERR_INVALID_ARG, "NULL argument");
  if (out_capacity < 0 || (out_capacity > 0 && !out)) return fail(PJ_ERR_INVALID_ARG, "bad output");
  if (!workspace || workspace_bytes < pj_join_workspace_size(total_bits))
    return fail(PJ_ERR_WORKSPACE, "join workspace too small");
  if (reinterpret_cast<uintptr_t>(workspace) % 256 != 0)
    return fail(PJ_ERR_INVALID_ARG, "workspace must be 256-byte aligned");
  const int cap = join_table_capacity(key_bytes) / 2 - 1;
  if (base < 1) base = 1;
  if (base > cap) base = cap;
  if (probe_slice <= 0) probe_slice = int64_t(base) * 8 < 16384 ? 16384 : int64_t(base) * 8;
  PJ_HIP(launch_join_generic(d_build_tuples, d_build_offsets, d_probe_tuples, d_probe_offsets, key_bytes, total_bits,
                             base, probe_slice, unique_build_keys, out, out_capacity, d_out_count,
                             static_cast<char*>(workspace), stream));
  return PJ_OK;
}

pj_status pj_hash_keys(const void* d_keys, int64_t n, int32_t key_bytes, uint32_t* d_out, hipStream_t stream) {
  if ((key_bytes != 4 && key_bytes != 8) || n < 0 || (n > 0 && (!d_keys || !d_out)))
    return fail(PJ_ERR_INVALID_ARG, "bad hash_keys arguments");
  if (n == 0) return PJ_OK;
  PJ_HIP(hash_keys(d_keys, n, key_bytes, d_out, stream));
  return PJ_OK;
}

pj_status pj_fastdiv(const int32_t* d_n, int64_t count, int32_t divisor, int32_t* d_q, int32_t* d_r,
                     hipStream_t stream) {
  if (divisor == 0) return fail(PJ_ERR_INVALID_ARG, "divisor must be non-zero");
  if (count < 0 || (count > 0 && !d_n)) return fail(PJ_ERR_INVALID_ARG, "bad fastdiv arguments");
  if (count == 0) return PJ_OK;
  PJ_HIP(fastdiv_run(d_n, count, divisor, d_q, d_r, stream));
  return PJ_OK;
}

// ---------------------------------------------------------------------------
// Generators
// ---------------------------------------------------------------------------
#define PJ_GEN_CHECK(ptr, n)                                                              \
  do {                                                                                    \
    if ((n) < 0 || ((n) > 0 && !(ptr))) return fail(PJ_ERR_INVALID_ARG, "bad generator output"); \
    if ((n) == 0) return PJ_OK;                                                           \
  } while (0)

pj_status pj_gen_build_unique_i32(int32_t* d_keys, int64_t n, int64_t domain, uint64_t perm_seed,
                                  hipStream_t stream) {
  PJ_GEN_CHECK(d_keys, n);
  if (domain < n || domain > (int64_t(1) << 32)) return fail(PJ_ERR_INVALID_ARG, "domain must be in [n, 2^32]");
  PJ_HIP(gen_build_unique_i32(d_keys, n, domain, perm_seed, stream));
  return PJ_OK;
}

pj_status pj_gen_build_dup_i32(int32_t* d_keys, int64_t n, int64_t m, int64_t domain, uint64_t perm_seed,
                               uint64_t seed, hipStream_t stream) {
  PJ_GEN_CHECK(d_keys, n);
  if (m < 1 || m > domain || domain > (int64_t(1) << 32)) return fail(PJ_ERR_INVALID_ARG, "need 1 <= m <= domain");
  PJ_HIP(gen_build_dup_i32(d_keys, n, m, domain, perm_seed, seed, stream));
  return PJ_OK;
}

pj_status pj_gen_probe_i32(int32_t* d_keys, int64_t n, int64_t domain, int64_t n_build, uint64_t perm_seed,
                           uint64_t seed, double sel, int64_t nonmatch_base, int64_t nonmatch_range,
                           hipStream_t stream) {
  PJ_GEN_CHECK(d_keys, n);
  if (n_build < 1 || n_build > domain || domain > (int64_t(1) << 32) || nonmatch_range < 0)
    return fail(PJ_ERR_INVALID_ARG, "bad probe generator sizes");
  if (nonmatch_range == 0 && domain == n_build && sel < 1.0)
    return fail(PJ_ERR_INVALID_ARG, "no non-matching keys left in the domain");
  PJ_HIP(gen_probe_i32(d_keys, n, domain, n_build, perm_seed, seed, sel, nonmatch_base, nonmatch_range, stream));
  return PJ_OK;
}

pj_status pj_zipf_lut(double* h_lut, int64_t n, double theta) {
  if (!h_lut || n < 1) return fail(PJ_ERR_INVALID_ARG, "bad zipf table arguments");
  double scale = 0.0;
  for (int64_t i = 1; i <= n; ++i) scale += 1.0 / std::pow(double(i), theta);
  double sum = 0.0;
  for (int64_t i = 1; i <= n; ++i) {
    sum += 1.0 / std::pow(double(i), theta);
    h_lut[i - 1] = sum / scale;
  }
  return PJ_OK;
}

pj_status pj_gen_probe_zipf_i32(int32_t* d_keys, int64_t count, const double* d_lut, int64_t n,
                                uint64_t rank_perm_seed, uint64_t seed, hipStream_t stream) {
  PJ_GEN_CHECK(d_keys, count);
  if (!d_lut || n < 1 || n > (int64_t(1) << 32)) return fail(PJ_ERR_INVALID_ARG, "bad zipf arguments");
  PJ_HIP(gen_probe_zipf_i32(d_keys, count, d_lut, n, rank_perm_seed, seed, stream));
  return PJ_OK;
}

pj_status pj_gen_build_i64(int64_t* d_keys, int64_t n, uint64_t key_seed, hipStream_t stream) {
  PJ_GEN_CHECK(d_keys, n);
  PJ_HIP(gen_build_i64(d_keys, n, key_seed, stream));
  return PJ_OK;
}

pj_status pj_gen_probe_i64(int64_t* d_keys, int64_t n, int64_t n_build, uint64_t key_seed, uint64_t seed, double sel,
                           hipStream_t stream) {
  PJ_GEN_CHECK(d_keys, n);
  if (n_build < 1) return fail(PJ_ERR_INVALID_ARG, "n_build must be >= 1");
  PJ_HIP(gen_probe_i64(d_keys, n, n_build, key_seed, seed, sel, stream));
  return PJ_OK;
}

pj_status pj_gen_probe_zipf_i64(int64_t* d_keys, int64_t count, const double* d_lut, int64_t n_build,
                                uint64_t key_seed, uint64_t rank_perm_seed, uint64_t seed, hipStream_t stream) {
  PJ_GEN_CHECK(d_keys, count);
  if (!d_lut || n_build < 1) return fail(PJ_ERR_INVALID_ARG, "bad zipf arguments");
  PJ_HIP(gen_probe_zipf_i64(d_keys, count, d_lut, n_build, key_seed, rank_perm_seed, seed, stream));
  return PJ_OK;
}

pj_status pj_gen_row_ids(int32_t* d_pay, int64_t n, hipStream_t stream) {
  PJ_GEN_CHECK(d_pay, n);
  PJ_HIP(gen_row_ids(d_pay, n, stream));
  return PJ_OK;
}

// ---------------------------------------------------------------------------
// Profiling
// ---------------------------------------------------------------------------
pj_status pj_profile_enable(int32_t on) {
  g_prof.on = on != 0;
  g_prof.recorded = false;
  g_prof.count = 0;
  return PJ_OK;
}

pj_status pj_profile_select(uint32_t marks) {
  if (marks == 0 || marks > 0xFFu) return fail(PJ_ERR_INVALID_ARG, "marks must be a non-empty subset of 0xFF");
  g_prof.marks = marks;
  return PJ_OK;
}

// The last event that set k recorded.
static hipEvent_t prof_last(int k) {
  int i = 7;
  while (i > 0 && !((g_prof.set_marks[k] >> i) & 1u)) --i;
  return g_prof.ev[k][i];
}

// Phase times of event set k, whose last event has completed: a phase whose
// two boundaries were not both recorded reads 0.
static pj_status prof_times(int k, pj_phase_times* t) {
  std::memset(t, 0, sizeof(*t));
  const hipEvent_t* ev = g_prof.ev[k];
  const uint32_t m = g_prof.set_marks[k];
  float ms[7] = {};
  for (int i = 0; i < 7; ++i)
    if (((m >> i) & 3u) == 3u) PJ_HIP(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
  t->histogram_ms = ms[0];
  t->offsets_ms = ms[1];
  t->pass1_ms = ms[2];
  t->pass2_ms = ms[3];
  t->skew_ms = ms[4];
  t->join_ms = ms[5];
  t->compact_ms = ms[6];
  if ((m & 0x81u) == 0x81u) PJ_HIP(hipEventElapsedTime(&t->total_ms, ev[0], ev[7]));
  t->valid = 1;
  return PJ_OK;
}

pj_status pj_profile_read(pj_phase_times* t) {
  if (!t) return fail(PJ_ERR_INVALID_ARG, "times is NULL");
  std::memset(t, 0, sizeof(*t));
  if (!g_prof.on || !g_prof.recorded) return PJ_OK;
  const int k = (g_prof.head + kProfRing - 1) % kProfRing;
  PJ_HIP(hipEventSynchronize(prof_last(k)));
  return prof_times(k, t);
}

pj_status pj_profile_read_recent(pj_phase_times* times, int32_t n, int32_t* got) {
  if (!times || !got) return fail(PJ_ERR_INVALID_ARG, "times or got is NULL");
  if (n < 0) return fail(PJ_ERR_INVALID_ARG, "n must be >= 0");
  *got = 0;
  if (!g_prof.on) return PJ_OK;
  const int m = n < g_prof.count ? n : g_prof.count;
  if (m == 0) return PJ_OK;
  const int newest = (g_prof.head + kProfRing - 1) % kProfRing;
  PJ_HIP(hipEventSynchronize(prof_last(newest)));
  for (int j = 0; j < m; ++j) {
    const pj_status st = prof_times((g_prof.head + kProfRing - m + j) % kProfRing, &times[j]);
    if (st != PJ_OK) return st;
  }
  *got = m;
  return PJ_OK;
}

// ---------------------------------------------------------------------------
// Learned planner
// ---------------------------------------------------------------------------
size_t pj_planner_param_count(void) { return kPlanParams; }

int32_t pj_join_chunk_capacity(int32_t key_bytes) {
  return key_bytes == 4 || key_bytes == 8 ? join_table_capacity(key_bytes) / 2 - 1 : 0;
}

size_t pj_planner_scratch_bytes(int32_t batch) {
  return batch < 1 ? 0 : planner_scratch_floats(batch) * sizeof(float);
}

pj_status pj_planner_sketch(const pj_table* build, const pj_table* probe, float* d_features, uint64_t* d_cand_keys,
                            uint32_t* d_cand_sample_counts, hipStream_t stream) {
  pj_status st = check_pair(build, probe);
  if (st != PJ_OK) return st;
  if (!d_features || !d_cand_keys || !d_cand_sample_counts) return fail(PJ_ERR_INVALID_ARG, "NULL output");
  PJ_HIP(planner_sketch(probe->keys, probe->n, probe->key_bytes, build->n, d_features, d_cand_keys,
                        d_cand_sample_counts, stream));
  return PJ_OK;
}

pj_status pj_planner_candidate_counts(const pj_table* probe, const uint64_t* d_cand_keys,
                                      const uint32_t* d_cand_sample_counts, uint64_t* d_counts, hipStream_t stream) {
  pj_status st = check_table(probe, "probe");
  if (st != PJ_OK) return st;
  if (!d_cand_keys || !d_cand_sample_counts || !d_counts) return fail(PJ_ERR_INVALID_ARG, "NULL argument");
  PJ_HIP(planner_count_candidates(probe->keys, probe->n, probe->key_bytes, d_cand_keys, d_cand_sample_counts,
                                  reinterpret_cast<unsigned long long*>(d_counts), stream));
  return PJ_OK;
}

pj_status pj_planner_forward(const float* d_weights, const float* d_features, int32_t batch, float* d_logits,
                             void* d_scratch, size_t scratch_bytes, hipStream_t stream) {
  if (!d_weights || !d_features || !d_logits || !d_scratch || batch < 1)
    return fail(PJ_ERR_INVALID_ARG, "bad planner forward arguments");
  if (scratch_bytes < pj_planner_scratch_bytes(batch))
    return fail(PJ_ERR_WORKSPACE, "planner scratch needs %zu bytes", pj_planner_scratch_bytes(batch));
  PJ_HIP(planner_forward(d_weights, d_features, batch, d_logits, static_cast<float*>(d_scratch), stream));
  return PJ_OK;
}

pj_status pj_planner_train_step(float* d_weights, float* d_adam_m, float* d_adam_v, int32_t step, float lr,
                                const float* d_features, const int32_t* d_labels, const float* d_heavy,
                                int32_t batch, float* d_loss, void* d_scratch, size_t scratch_bytes,
                                hipStream_t stream) {
  if (!d_weights || !d_adam_m || !d_adam_v || !d_features || !d_labels || !d_heavy || !d_loss || !d_scratch ||
      batch < 1 || step < 1 || !(lr > 0.f))
    return fail(PJ_ERR_INVALID_ARG, "bad planner training arguments");
  if (scratch_bytes < pj_planner_scratch_bytes(batch))
    return fail(PJ_ERR_WORKSPACE, "planner scratch needs %zu bytes", pj_planner_scratch_bytes(batch));
  PJ_HIP(planner_train_step(d_weights, d_adam_m, d_adam_v, step, lr, d_features, d_labels, d_heavy, batch, d_loss,
                            static_cast<float*>(d_scratch), stream));
  return PJ_OK;
}

size_t pj_plan_learned_workspace_size(void) {
  Carver c{nullptr};
  c.take(kPlanFeat * sizeof(float));
  c.take(kPlanCand * 8);
  c.take(kPlanCand * 4);
  c.take(kPlanCand * 8);
  c.take(kPlanOut * sizeof(float));
  c.take(planner_scratch_floats(1) * sizeof(float));
  return c.off;
}

}  // extern "C"

namespace pj {

// The plan the logits predict (bits and pass count; the rest as the plain
// plan), and whether it may be taken at all: its build chunks must fit the
// LDS table.
static bool learned_plan_of(const float* logits, const pj_table* build, pj_plan* learned) {
  auto argmax = [&](int off, int n) {
    int best = 0;
    for (int c = 1; c < n; ++c)
      if (logits[off + c] > logits[off + best]) best = c;
    return best;
  };
  plain_plan(build->n, build->key_bytes, learned);
  const int passes = 1 + argmax(2 * kPlanBitsCls, kPlanPassCls);
  const int b1 = 6 + argmax(0, kPlanBitsCls), b2 = 6 + argmax(kPlanBitsCls, kPlanBitsCls);
  if (passes == 1) {
    learned->passes = 1;
    learned->bits1 = b1;
    learned->bits2 = 0;
  } else {  // three passes are not implemented: two passes with the predicted bits
    learned->passes = 2;
    learned->bits1 = b1;
    learned->bits2 = b1 + b2 > kMaxCombinedBits ? kMaxCombinedBits - b1 : b2;
  }
  learned->source = 1;
  // The plan fits when its partitions are no larger than the plain plan's aim
  // or its chunks (at the plan's factor) fit the table.
  const double cap = double(join_table_capacity(build->key_bytes) / 2 - 1);
  const double avg = double(build->n) / double(int64_t(1) << (learned->bits1 + learned->bits2));
  const bool fits = avg <= partition_target(build->key_bytes) || learned->max_partition_size_factor * avg <= cap;
  return fits && check_plan(learned) == PJ_OK;
}

static bool same_plan(const pj_plan& a, const pj_plan& b) {
  return a.passes == b.passes && a.bits1 == b.bits1 && a.bits2 == b.bits2 &&
         a.max_partition_size_factor == b.max_partition_size_factor && a.probe_slice == b.probe_slice &&
         a.unique_build_keys == b.unique_build_keys;
}

}  // namespace pj

extern "C" {

pj_status pj_plan_decode(const float* logits, const uint32_t* cand_sample_counts, const uint64_t* cand_exact_counts,
                         const pj_table* build, const pj_table* probe, pj_plan* plan, int32_t* fallback) {
  pj_status st = check_pair(build, probe, false);
  if (st != PJ_OK) return st;
  if (!logits || !cand_sample_counts || !cand_exact_counts || !plan) return fail(PJ_ERR_INVALID_ARG, "NULL argument");
  // Heavy key: at least T = |S| / 1024 probe rows.  The predicted set
  // disagrees with the exact counts when a key predicted heavy has fewer than
  // T / 2 rows or a key not predicted heavy has 2T or more; keys within a
  // factor of two of T (where a 4096-key sample cannot decide) count either
  // way.  Any disagreement selects the plain plan.
  bool agree = true;
  const int hoff = 2 * kPlanBitsCls + kPlanPassCls;
  const uint64_t ns = uint64_t(probe->n);
  for (int j = 0; j < kPlanCand; ++j) {
    if (cand_sample_counts[j] == 0) continue;
    const bool pred = logits[hoff + j] > 0.f;
    const uint64_t c = cand_exact_counts[j];
    if (pred && c * 2048ull < ns) agree = false;         // predicted heavy, below T / 2
    if (!pred && c * 512ull >= ns && ns > 0) agree = false;  // missed, at least 2T
  }
  pj_plan learned;
  const bool fits = learned_plan_of(logits, build, &learned);
  if (agree && fits) {
    *plan = learned;
    if (fallback) *fallback = 0;
  } else {
    plain_plan(build->n, build->key_bytes, plan);
    if (fallback) *fallback = 1;
  }
  return PJ_OK;
}

pj_status pj_plan_learned(const pj_table* build, const pj_table* probe, const float* d_weights, pj_plan* plan,
                          int32_t* fallback, void* workspace, size_t workspace_bytes, hipStream_t stream) {
  pj_status st = check_pair(build, probe);
  if (st != PJ_OK) return st;
  if (!d_weights || !plan || !workspace) return fail(PJ_ERR_INVALID_ARG, "NULL argument");
  if (workspace_bytes < pj_plan_learned_workspace_size())
    return fail(PJ_ERR_WORKSPACE, "learned planning needs %zu bytes", pj_plan_learned_workspace_size());
  Carver c{static_cast<char*>(workspace)};
  float* feat = static_cast<float*>(c.take(kPlanFeat * sizeof(float)));
  uint64_t* cand = static_cast<uint64_t*>(c.take(kPlanCand * 8));
  uint32_t* cand_n = static_cast<uint32_t*>(c.take(kPlanCand * 4));
  unsigned long long* exact = static_cast<unsigned long long*>(c.take(kPlanCand * 8));
  float* logits = static_cast<float*>(c.take(kPlanOut * sizeof(float)));
  float* scratch = static_cast<float*>(c.take(planner_scratch_floats(1) * sizeof(float)));
  PJ_HIP(planner_sketch(probe->keys, probe->n, probe->key_bytes, build->n, feat, cand, cand_n, stream));
  PJ_HIP(planner_forward(d_weights, feat, 1, logits, scratch, stream));
  float h_logits[kPlanOut];
  PJ_HIP(hipMemcpyAsync(h_logits, logits, sizeof(h_logits), hipMemcpyDeviceToHost, stream));
  PJ_HIP(hipStreamSynchronize(stream));
  // The heavy-set check chooses between the learned and the plain plan; when
  // the model predicts the plain plan itself there is nothing to choose, and
  // the exact candidate counts (a read of the whole probe column) are skipped.
  pj_plan learned, plain;
  const bool fits = learned_plan_of(h_logits, build, &learned);
  plain_plan(build->n, build->key_bytes, &plain);
  if (fits && same_plan(learned, plain)) {
    *plan = learned;
    if (fallback) *fallback = 0;
    return PJ_OK;
  }
  PJ_HIP(planner_count_candidates(probe->keys, probe->n, probe->key_bytes, cand, cand_n, exact, stream));
  // cand_n and exact lie in that order in the workspace: one copy
  const char* lo = reinterpret_cast<const char*>(cand_n);
  const size_t span = size_t(reinterpret_cast<const char*>(exact + kPlanCand) - lo);
  alignas(16) char h_span[((kPlanCand * 4 + 255) & ~255) + kPlanCand * 8];
  if (span != sizeof(h_span)) return fail(PJ_ERR_UNSUPPORTED, "planner workspace layout changed");
  PJ_HIP(hipMemcpyAsync(h_span, lo, span, hipMemcpyDeviceToHost, stream));
  PJ_HIP(hipStreamSynchronize(stream));
  const uint32_t* h_cand_n = reinterpret_cast<const uint32_t*>(h_span);
  const uint64_t* h_exact = reinterpret_cast<const uint64_t*>(h_span + (reinterpret_cast<const char*>(exact) - lo));
  return pj_plan_decode(h_logits, h_cand_n, h_exact, build, probe, plan, fallback);
}

// ---------------------------------------------------------------------------
// Batched joins: one graph of many independent join pipelines
// ---------------------------------------------------------------------------
}  // extern "C"

struct pj_batch {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int32_t count = 0;
  int32_t chains = 0;
};

namespace pj {

// Longest-first assignment of the joins to `chains` chains by |R| + |S| (plus
// a fixed cost per join for its launches), so the chains finish together.
static std::vector<int> assign_chains(const pj_table* builds, const pj_table* probes, int count, int chains) {
  std::vector<int> order(count), chain_of(count);
  std::iota(order.begin(), order.end(), 0);
  auto cost = [&](int j) { return double(builds[j].n) + double(probes[j].n) + 2.0e5; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
  std::vector<double> load(chains, 0.0);
  for (int j : order) {
    const int c = int(std::min_element(load.begin(), load.end()) - load.begin());
    chain_of[j] = c;
    load[c] += cost(j);
  }
  return chain_of;
}

// Workspace of each chain (the largest of its joins), 256-byte aligned.
static pj_status chain_workspaces(const pj_table* builds, const pj_table* probes, const pj_plan* plans, int count,
                                  int chains, const std::vector<int>& chain_of, std::vector<size_t>* off,
                                  size_t* total) {
  std::vector<size_t> need(chains, 0);
  for (int j = 0; j < count; ++j) {
    size_t b = 0;
    pj_status st = pj_workspace_size(&builds[j], &probes[j], plans ? &plans[j] : nullptr, &b);
    if (st != PJ_OK) return st;
    need[chain_of[j]] = std::max(need[chain_of[j]], (b + 255) & ~size_t(255));
  }
  off->assign(chains, 0);
  size_t acc = 0;
  for (int c = 0; c < chains; ++c) {
    (*off)[c] = acc;
    acc += need[c];
  }
  *total = acc;
  return PJ_OK;
}

static pj_status check_batch(const pj_table* builds, const pj_table* probes, int32_t count, int32_t chains) {
  if (!builds || !probes || count < 1) return fail(PJ_ERR_INVALID_ARG, "a batch needs count >= 1 tables");
  if (chains < 1 || chains > 64) return fail(PJ_ERR_INVALID_ARG, "chains must be in [1, 64] (got %d)", chains);
  for (int j = 0; j < count; ++j) {
    pj_status st = check_pair(&builds[j], &probes[j], false);
    if (st != PJ_OK) return fail(st, "join %d: %s", j, g_err.c_str());
  }
  return PJ_OK;
}

}  // namespace pj

extern "C" {

pj_status pj_batch_workspace_size(const pj_table* builds, const pj_table* probes, const pj_plan* plans,
                                  int32_t count, int32_t chains, size_t* bytes) {
  pj_status st = check_batch(builds, probes, count, chains);
  if (st != PJ_OK) return st;
  if (!bytes) return fail(PJ_ERR_INVALID_ARG, "bytes is NULL");
  const std::vector<int> chain_of = assign_chains(builds, probes, count, chains);
  std::vector<size_t> off;
  return chain_workspaces(builds, probes, plans, count, chains, chain_of, &off, bytes);
}

pj_status pj_batch_create(const pj_table* builds, const pj_table* probes, const pj_plan* plans, int32_t count,
                          int32_t chains, pj_pair* out, const int64_t* out_offsets, const int64_t* out_caps,
                          int64_t* d_counts, void* workspace, size_t workspace_bytes, pj_batch** batch) {
  pj_status st = check_batch(builds, probes, count, chains);
  if (st != PJ_OK) return st;
  if (!out_offsets || !out_caps || !d_counts || !batch || !workspace)
    return fail(PJ_ERR_INVALID_ARG, "NULL batch argument");
  *batch = nullptr;
  for (int j = 0; j < count; ++j) {
    st = check_pair(&builds[j], &probes[j]);
    if (st != PJ_OK) return fail(st, "join %d: %s", j, g_err.c_str());
    if (out_offsets[j] < 0 || out_caps[j] < 0 || (out_caps[j] > 0 && !out))
      return fail(PJ_ERR_INVALID_ARG, "join %d: bad output offset or capacity", j);
  }
  const std::vector<int> chain_of = assign_chains(builds, probes, count, chains);
  std::vector<size_t> off;
  size_t need = 0;
  st = chain_workspaces(builds, probes, plans, count, chains, chain_of, &off, &need);
  if (st != PJ_OK) return st;
  if (workspace_bytes < need)
    return fail(PJ_ERR_WORKSPACE, "batch workspace of %zu bytes is smaller than the %zu required", workspace_bytes,
                need);
  if (reinterpret_cast<uintptr_t>(workspace) % 256 != 0)
    return fail(PJ_ERR_INVALID_ARG, "workspace must be 256-byte aligned");

  // Capture: a fork to one stream per chain, each join's pipeline in its
  // chain, a join back to the capture stream.  Profiling events stay off.
  const bool prof = g_prof.on;
  g_prof.on = false;
  hipStream_t cs = nullptr;
  std::vector<hipStream_t> streams(chains, nullptr);
  std::vector<hipEvent_t> done(chains, nullptr);
  hipEvent_t fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  pj_status result = PJ_OK;
  bool capturing = false;
  auto hip_ok = [&](hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    result = fail(PJ_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
    return false;
  };
  do {
    if (!hip_ok(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreateWithFlags")) break;
    bool ok = true;
    for (int c = 0; c < chains && ok; ++c)
      ok = hip_ok(hipStreamCreateWithFlags(&streams[c], hipStreamNonBlocking), "hipStreamCreateWithFlags") &&
           hip_ok(hipEventCreateWithFlags(&done[c], hipEventDisableTiming), "hipEventCreateWithFlags");
    if (!ok || !hip_ok(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "hipEventCreateWithFlags")) break;
    if (!hip_ok(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture")) break;
    capturing = true;
    if (!hip_ok(hipEventRecord(fork, cs), "hipEventRecord")) break;
    for (int c = 0; c < chains && ok; ++c) ok = hip_ok(hipStreamWaitEvent(streams[c], fork, 0), "hipStreamWaitEvent");
    if (!ok) break;
    for (int j = 0; j < count && ok; ++j) {
      const int c = chain_of[j];
      char* ws = static_cast<char*>(workspace) + off[c];
      const size_t ws_bytes = (c + 1 < chains ? off[c + 1] : need) - off[c];
      const pj_status js = run_join(&builds[j], &probes[j], plans ? &plans[j] : nullptr, out + out_offsets[j],
                                    out_caps[j], d_counts + j, ws, ws_bytes, streams[c]);
      if (js != PJ_OK) {
        result = fail(js, "join %d: %s", j, g_err.c_str());
        ok = false;
      }
    }
    if (!ok) break;
    for (int c = 0; c < chains && ok; ++c)
      ok = hip_ok(hipEventRecord(done[c], streams[c]), "hipEventRecord") &&
           hip_ok(hipStreamWaitEvent(cs, done[c], 0), "hipStreamWaitEvent");
    if (!ok) break;
    capturing = false;
    if (!hip_ok(hipStreamEndCapture(cs, &graph), "hipStreamEndCapture")) break;
    if (!hip_ok(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate")) break;
  } while (false);
  if (capturing) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(cs, &g);
    if (g) (void)hipGraphDestroy(g);
  }
  g_prof.on = prof;
  for (int c = 0; c < chains; ++c) {
    if (done[c]) (void)hipEventDestroy(done[c]);
    if (streams[c]) (void)hipStreamDestroy(streams[c]);
  }
  if (fork) (void)hipEventDestroy(fork);
  if (cs) (void)hipStreamDestroy(cs);
  if (result != PJ_OK) {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    return result;
  }
  pj_batch* b = new pj_batch;
  b->graph = graph;
  b->exec = exec;
  b->count = count;
  b->chains = chains;
  *batch = b;
  return PJ_OK;
}

pj_status pj_batch_launch(pj_batch* batch, hipStream_t stream) {
  if (!batch || !batch->exec) return fail(PJ_ERR_INVALID_ARG, "batch is NULL");
  PJ_HIP(hipGraphLaunch(batch->exec, stream));
  return PJ_OK;
}

pj_status pj_batch_destroy(pj_batch* batch) {
  if (!batch) return PJ_OK;
  if (batch->exec) (void)hipGraphExecDestroy(batch->exec);
  if (batch->graph) (void)hipGraphDestroy(batch->graph);
  delete batch;
  return PJ_OK;
}

}  // extern "C"
