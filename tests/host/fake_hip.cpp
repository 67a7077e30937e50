// fake_hip.cpp -- a stand-in for the HIP runtime, for running libpj.so's host
// code on a machine without a GPU (tests/host: an AddressSanitizer build of the
// host side of every source file, kernels omitted).
//
// "Device memory" is host memory that the driver registers (fake_register);
// copies are memcpy; streams, events and graphs are bookkeeping.  Kernels do
// not run.  Every launch is checked instead:
//   - grid and block within the hardware limits, dynamic LDS <= 160 KB;
//   - stream capture: a launch on a capturing stream joins the capture, and a
//     graph replay checks its recorded launches again;
//   - for the join pipeline's kernels, every buffer argument lies inside one
//     registered allocation, over the extent the kernel can touch (derived
//     from its other arguments the way the kernel indexes it).
// Violations are counted and printed; fake_errors() returns the count.
#include <cxxabi.h>
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../partitioned_join_amd/csrc/pj_internal.h"

namespace {

struct Alloc {
  uintptr_t lo, hi;
  std::string name;
};
// Registered allocations; __device__ variables are registered by the objects'
// constructors, which may run before this file's, hence a function-local static.
std::vector<Alloc>& allocs() {
  static std::vector<Alloc> v;
  return v;
}
int g_errors = 0;
long g_launches = 0;
// Kernel names by host stub; filled by the objects' constructors, which may
// run before this file's, hence a function-local static.
std::map<const void*, std::string>& funcs() {
  static std::map<const void*, std::string> m;
  return m;
}

void err(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "[fake_hip] ERROR: ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  va_end(ap);
  ++g_errors;
}

// [p, p + bytes) must lie inside one registered allocation.
void need(const char* kern, const char* what, const void* p, uint64_t bytes) {
  if (bytes == 0) return;
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  for (const Alloc& al : allocs())
    if (a >= al.lo && a < al.hi) {
      if (a + bytes > al.hi)
        err("%s: %s [%p, +%llu) overruns %s by %llu bytes", kern, what, p, (unsigned long long)bytes,
            al.name.c_str(), (unsigned long long)(a + bytes - al.hi));
      return;
    }
  err("%s: %s %p (+%llu bytes) is in no registered allocation", kern, what, p, (unsigned long long)bytes);
}

struct CallCfg {
  dim3 grid, block;
  size_t shmem;
  hipStream_t stream;
};
thread_local std::vector<CallCfg> g_cfg;

struct Launch {
  const void* fn;
  dim3 grid, block;
  size_t shmem;
  std::vector<char> args;  // the first argument's bytes (all kernels checked take a struct or pointers)
  std::vector<uintptr_t> words;
};

struct Capture;
struct Stream {
  std::shared_ptr<Capture> cap;
};
struct Event {
  std::shared_ptr<Capture> cap;
};
struct Capture {
  std::vector<Launch> launches;
  bool open = true;
};
struct Graph {
  std::vector<Launch> launches;
};
struct Exec {
  std::vector<Launch> launches;
};

Stream g_null_stream;
Stream* S(hipStream_t s) { return s ? reinterpret_cast<Stream*>(s) : &g_null_stream; }

std::string demangle(const char* name) {
  int st = 0;
  char* d = abi::__cxa_demangle(name, nullptr, nullptr, &st);
  std::string r = st == 0 && d ? d : name;
  free(d);
  if (r.compare(0, 5, "void ") == 0) r = r.substr(5);  // function templates carry their return type
  return r;
}

bool starts(const std::string& s, const char* p) { return s.compare(0, strlen(p), p) == 0; }

int tmpl_int(const std::string& name, int index) {
  // index-th integer template argument of "pj::k_x<a, b, ...>"
  const size_t lt = name.find('<');
  if (lt == std::string::npos) return -1;
  const char* c = name.c_str() + lt + 1;
  for (int i = 0; i < index; ++i) {
    c = strchr(c, ',');
    if (!c) return -1;
    ++c;
  }
  return atoi(c);
}

using namespace pj;

void check_table(const char* k, const PartTable& T, const PartParams& pp, int kb, bool scatter1, bool scatter2) {
  const uint64_t tb = kb == 4 ? 8 : 16;
  const bool deep = pp.passes == 2 && pp.B > kMaxCombinedBits;
  if (scatter1 || !scatter2) need(k, "keys", T.keys, uint64_t(T.n) * kb);
  if (scatter1 && T.pay) need(k, "payload", T.pay, uint64_t(T.n) * 4);
  need(k, "mat", T.mat, uint64_t(T.nseg) * (deep ? pp.F1 : pp.P) * 4);
  if (pp.passes == 2 && !deep) need(k, "matc", T.matc, uint64_t(T.nseg) * pp.F1 * 4);
  if (deep) {
    need(k, "offc", T.offc, uint64_t(pp.F1 + 1) * 4);
    need(k, "h2", T.h2, deep_item_capacity(T.n, kb, pp) * pp.F2 * 4);
  }
  if (scatter1 || scatter2) {
    need(k, "off", T.off, uint64_t(pp.P + 1) * 4);
    need(k, "buf1", T.buf1, uint64_t(T.n) * tb);
  }
  if (scatter2) need(k, "buf2", T.buf2, uint64_t(T.n) * tb);
  if (T.nseg % 8 != 0 || T.nseg == 0) err("%s: nseg %u is not a positive multiple of 8", k, T.nseg);
}

void check_join_args(const char* k, const JoinArgs& A, uint32_t grid) {
  const int kb = A.bucket_slots == 4 ? 4 : 8;
  const uint64_t tb = kb == 4 ? 8 : 16;
  need(k, "roff", A.roff, uint64_t(A.P + 1) * 4);
  need(k, "soff", A.soff, uint64_t(A.P + 1) * 4);
  need(k, "item_scan", A.item_scan, uint64_t(A.P + 1) * 8);
  if (A.nr != 0xFFFFFFFFu) need(k, "build tuples", A.rt, uint64_t(A.nr) * tb);
  if (A.ns != 0xFFFFFFFFu) need(k, "probe tuples", A.st, uint64_t(A.ns) * tb);
  need(k, "out", A.out, uint64_t(A.capacity) * 8);
  need(k, "spill", A.spill, uint64_t(grid) * uint64_t(kOutChunk) * 8);
  need(k, "chunk_ctr", A.chunk_ctr, 4);
  need(k, "wg_last", A.wg_last, uint64_t(grid) * 8);
  need(k, "desc", A.desc, A.desc_cap * 32);
  const int table_cap = pj::join_table_capacity(kb);
  if (A.base < 1 || A.base > table_cap / 2 - 1) err("%s: base %d outside [1, %d]", k, A.base, table_cap / 2 - 1);
  // power-of-two bucket tables: within the LDS slots, and more slots than the largest chunk
  const uint32_t cb = A.cap_buckets, bs = uint32_t(A.bucket_slots);
  if (cb == 0 || (cb & (cb - 1)) != 0 || cb * bs > uint32_t(table_cap) || cb * bs <= uint32_t(A.base))
    err("%s: cap_buckets %u (x %u slots) against table_cap %d and base %d", k, cb, bs, table_cap, A.base);
  if (A.slice < 1) err("%s: probe slice %lld < 1", k, (long long)A.slice);
}

void check_launch(const Launch& L) {
  auto it = funcs().find(L.fn);
  const std::string name = it == funcs().end() ? std::string("?") : it->second;
  const char* k = name.c_str();
  ++g_launches;
  if (L.grid.x < 1 || L.grid.y < 1 || L.grid.z < 1 || L.grid.x > 0x7fffffffu || L.grid.y > 65535u)
    err("%s: grid %u x %u x %u", k, L.grid.x, L.grid.y, L.grid.z);
  const uint64_t threads = uint64_t(L.block.x) * L.block.y * L.block.z;
  if (threads < 1 || threads > 1024) err("%s: %llu threads per block", k, (unsigned long long)threads);
  if (L.shmem > 160 * 1024) err("%s: %zu bytes of dynamic LDS", k, L.shmem);
  const void* a0 = L.args.data();
  const uintptr_t* w = L.words.data();
  if (starts(name, "pj::k_fill_u32")) {
    need(k, "p", reinterpret_cast<void*>(w[0]), uint64_t(int64_t(w[2])) * 4);
  } else if (starts(name, "pj::k_fill_u64")) {
    need(k, "p", reinterpret_cast<void*>(w[0]), uint64_t(int64_t(w[2])) * 8);
  } else if (starts(name, "pj::k_hist<")) {
    const PartTable* t = static_cast<const PartTable*>(a0);
    const PartParams& pp = *reinterpret_cast<const PartParams*>(reinterpret_cast<const char*>(a0) + 2 * sizeof(PartTable));
    for (uint32_t y = 0; y < L.grid.y; ++y) {
      check_table(k, t[y], pp, tmpl_int(name, 0), false, false);
      if (t[y].nseg > L.grid.x) err("%s: grid %u < nseg %u", k, L.grid.x, t[y].nseg);
    }
  } else if (starts(name, "pj::k_pass1<") || starts(name, "pj::k_pass2<")) {
    const bool two = starts(name, "pj::k_pass2<");
    const PartTable* t = static_cast<const PartTable*>(a0);
    const PartParams& pp = *reinterpret_cast<const PartParams*>(reinterpret_cast<const char*>(a0) + 2 * sizeof(PartTable));
    const int ntab = two ? 2 : int(L.grid.y);
    for (int y = 0; y < ntab; ++y) {
      if (two && t[y].n == 0) continue;
      check_table(k, t[y], pp, tmpl_int(name, 0), !two, two);
      if (!two && t[y].nseg > L.grid.x) err("%s: grid %u < nseg %u", k, L.grid.x, t[y].nseg);
    }
  } else if (starts(name, "pj::k_hist2<")) {
    const PartTable* t = static_cast<const PartTable*>(a0);
    const PartParams& pp = *reinterpret_cast<const PartParams*>(reinterpret_cast<const char*>(a0) + 2 * sizeof(PartTable));
    const int kb = tmpl_int(name, 0);
    for (int y = 0; y < 2; ++y) {
      if (t[y].n == 0) continue;
      need(k, "buf1", t[y].buf1, uint64_t(t[y].n) * (kb == 4 ? 8 : 16));
      need(k, "offc", t[y].offc, uint64_t(pp.F1 + 1) * 4);
      need(k, "h2", t[y].h2, deep_item_capacity(t[y].n, kb, pp) * pp.F2 * 4);
      need(k, "ipre", t[y].ipre, uint64_t((pp.F1 + 7) / 8 * 8 + 1) * 4);
    }
  } else if (starts(name, "pj::k_segscan")) {
    const PartTable* t = static_cast<const PartTable*>(a0);
    const PartParams& pp = *reinterpret_cast<const PartParams*>(reinterpret_cast<const char*>(a0) + 2 * sizeof(PartTable));
    for (uint32_t y = 0; y < L.grid.y; ++y) {
      need(k, "tot", t[y].tot, uint64_t(pp.P) * 4);
      need(k, "ipre", t[y].ipre, uint64_t((pp.F1 + 7) / 8 * 8 + 1) * 4);
    }
  } else if (starts(name, "pj::k_items_expand")) {
    const JoinArgs& A = *static_cast<const JoinArgs*>(a0);
    need(k, "item_scan", A.item_scan, uint64_t(A.P + 1) * 8);
    need(k, "roff", A.roff, uint64_t(A.P + 1) * 4);
    need(k, "soff", A.soff, uint64_t(A.P + 1) * 4);
  } else if (starts(name, "pj::k_join<")) {
    check_join_args(k, *static_cast<const JoinArgs*>(a0), L.grid.x);
    // keys, payloads and a claim counter per bucket, plus the stage; two workgroups share a CU
    const JoinArgs& A = *static_cast<const JoinArgs*>(a0);
    const uint64_t kb = A.bucket_slots == 4 ? 4 : 8;
    const uint64_t stage = uint64_t(tmpl_int(name, 3)) * uint64_t(L.block.x) * 8;
    const uint64_t want = uint64_t(A.cap_buckets) * (uint64_t(A.bucket_slots) * (kb + 4) + 4) + stage;
    if (L.shmem < want) err("%s: %zu bytes of LDS, the table and stage need %llu", k, L.shmem, (unsigned long long)want);
    if (L.shmem > 80 * 1024) err("%s: %zu bytes of LDS leave no room for a second workgroup per CU", k, L.shmem);
  }
}

}  // namespace

extern "C" {

// ---- driver-side hooks ----
void fake_register(const void* p, size_t bytes, const char* name) {
  allocs().push_back({reinterpret_cast<uintptr_t>(p), reinterpret_cast<uintptr_t>(p) + bytes, name});
}
void fake_unregister(const void* p) {
  std::vector<Alloc>& v = allocs();
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i].lo == reinterpret_cast<uintptr_t>(p)) {
      v.erase(v.begin() + long(i));
      return;
    }
}
int fake_errors(void) { return g_errors; }
long fake_launches(void) { return g_launches; }

// ---- registration (called by each object's constructor) ----
void** __hipRegisterFatBinary(const void*) {
  static void* mod[4];
  return mod;
}
void __hipUnregisterFatBinary(void**) {}
void __hipRegisterFunction(void**, const void* host_fn, char*, const char* device_name, unsigned int, uint3*, uint3*,
                           dim3*, dim3*, int*) {
  funcs()[host_fn] = demangle(device_name);
}
void __hipRegisterVar(void**, void* var, char*, char* name, int, size_t size, int, int) {
  allocs().push_back({reinterpret_cast<uintptr_t>(var), reinterpret_cast<uintptr_t>(var) + size, name});
}

hipError_t __hipPushCallConfiguration(dim3 grid, dim3 block, size_t shmem, hipStream_t stream) {
  g_cfg.push_back({grid, block, shmem, stream});
  return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3* grid, dim3* block, size_t* shmem, hipStream_t* stream) {
  if (g_cfg.empty()) return hipErrorInvalidConfiguration;
  const CallCfg c = g_cfg.back();
  g_cfg.pop_back();
  *grid = c.grid;
  *block = c.block;
  *shmem = c.shmem;
  *stream = c.stream;
  return hipSuccess;
}

hipError_t hipLaunchKernel(const void* fn, dim3 grid, dim3 block, void** args, size_t shmem, hipStream_t stream) {
  Launch L{fn, grid, block, shmem, {}, {}};
  // The first argument's bytes (structs up to 1 KB) and the first 12 arguments
  // read as words (pointer and integer arguments of the plain kernels).
  auto it = funcs().find(fn);
  const std::string name = it == funcs().end() ? std::string("?") : it->second;
  size_t first = 8;
  if (starts(name, "pj::k_hist<") || starts(name, "pj::k_pass1<") || starts(name, "pj::k_pass2<") ||
      starts(name, "pj::k_hist2<") || starts(name, "pj::k_segscan"))
    first = 2 * sizeof(PartTable) + sizeof(PartParams);  // t0, t1, pp: separate arguments, copied one by one
  else if (starts(name, "pj::k_items_expand") || starts(name, "pj::k_join<"))
    first = sizeof(JoinArgs);
  L.args.resize(first);
  if (first == 2 * sizeof(PartTable) + sizeof(PartParams)) {
    memcpy(L.args.data(), args[0], sizeof(PartTable));
    memcpy(L.args.data() + sizeof(PartTable), args[1], sizeof(PartTable));
    memcpy(L.args.data() + 2 * sizeof(PartTable), args[2], sizeof(PartParams));
  } else {
    memcpy(L.args.data(), args[0], first);
  }
  if (starts(name, "pj::k_fill_u")) L.words = {*static_cast<uintptr_t*>(args[0]), 0, uintptr_t(*static_cast<int64_t*>(args[2]))};
  check_launch(L);
  Stream* s = S(stream);
  if (s->cap && s->cap->open) s->cap->launches.push_back(L);
  return hipSuccess;
}

hipError_t hipGetLastError(void) { return hipSuccess; }
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
// A __device__ variable's host-side stand-in is the "device" copy here; it is
// registered as an allocation of its size (given by the driver via
// fake_symbol_size, since the registration call is not informative here).
hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
  *p = const_cast<void*>(symbol);
  return hipSuccess;
}
const char* hipGetErrorString(hipError_t) { return "fake_hip error"; }

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  *s = reinterpret_cast<hipStream_t>(new Stream);
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  if (!s) return hipErrorInvalidHandle;
  Stream* st = S(s);
  if (st->cap && st->cap->open) err("stream destroyed while capturing");
  delete st;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  if (S(s)->cap && S(s)->cap->open) {
    err("hipStreamSynchronize on a capturing stream");
    return hipErrorStreamCaptureUnsupported;
  }
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = reinterpret_cast<hipEvent_t>(new Event);
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned int) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  if (!e) return hipErrorInvalidHandle;
  delete reinterpret_cast<Event*>(e);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  if (!e) return hipErrorInvalidHandle;
  reinterpret_cast<Event*>(e)->cap = S(s)->cap && S(s)->cap->open ? S(s)->cap : nullptr;
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 0.f;
  return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned int) {
  if (!e) return hipErrorInvalidHandle;
  Event* ev = reinterpret_cast<Event*>(e);
  if (ev->cap) S(s)->cap = ev->cap;  // the stream joins the capture
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (!s) return hipErrorInvalidValue;
  S(s)->cap = std::make_shared<Capture>();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  Stream* st = S(s);
  if (!st->cap || !st->cap->open) return hipErrorIllegalState;
  Graph* gr = new Graph;
  gr->launches = st->cap->launches;
  st->cap->open = false;
  st->cap.reset();
  *g = reinterpret_cast<hipGraph_t>(gr);
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  Exec* e = new Exec;
  e->launches = reinterpret_cast<Graph*>(g)->launches;
  *x = reinterpret_cast<hipGraphExec_t>(e);
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t) {
  if (!x) return hipErrorInvalidHandle;
  for (const Launch& L : reinterpret_cast<Exec*>(x)->launches) check_launch(L);
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
  delete reinterpret_cast<Exec*>(x);
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete reinterpret_cast<Graph*>(g);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t s) {
  if (S(s)->cap && S(s)->cap->open) err("hipMemcpyAsync on a capturing stream");
  memcpy(dst, src, bytes);
  return hipSuccess;
}

}  // extern "C"
