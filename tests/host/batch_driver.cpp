// batch_driver.cpp -- drives libpj.so's host code (built with
// AddressSanitizer against fake_hip.cpp, no GPU) through a script of C-ABI
// calls: eager joins, and batched joins created, launched and destroyed in a
// given order with given per-join plans.  Exit status 0 when AddressSanitizer
// found nothing, no call failed unexpectedly and fake_hip recorded no launch
// whose arguments reach outside their buffers.
//
// Script (stdin), one command per line:
//   joins N                     then N lines "n_build n_probe key_bytes"
//   eager                       pj_partitioned_join of every join, default plans
//   eager_short                 the same, but the workspace registered with the
//                               stand-in runtime is 4 KB shorter than the size
//                               passed (a planted mistake the checks must catch)
//   noalloc                     pj_partitioned_join_noalloc of every join
//   create ID CHAINS null|plans then, for "plans", N lines
//                               "passes bits1 bits2 factor unique slice source"
//   launch ID REPS
//   destroy ID
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/pj.h"

extern "C" {
void fake_register(const void* p, size_t bytes, const char* name);
void fake_unregister(const void* p);
int fake_errors(void);
long fake_launches(void);
}

namespace {

int g_fail = 0;

void* dev_alloc(size_t bytes, const char* name) {
  if (bytes == 0) return nullptr;
  const size_t r = (bytes + 255) & ~size_t(255);
  void* p = aligned_alloc(256, r);
  memset(p, 0, r);
  fake_register(p, bytes, name);
  return p;
}
void dev_free(void* p) {
  if (!p) return;
  fake_unregister(p);
  free(p);
}

struct Batch {
  pj_batch* h = nullptr;
  void* out = nullptr;
  void* counts = nullptr;
  void* ws = nullptr;
  std::vector<int64_t> offs, caps;
  std::vector<pj_plan> plans;
};

void expect_ok(pj_status st, const char* what) {
  if (st != PJ_OK) {
    fprintf(stderr, "[driver] %s failed: status %d: %s\n", what, int(st), pj_last_error());
    ++g_fail;
  }
}

}  // namespace

int main() {
  std::vector<pj_table> tb, tp;
  std::vector<void*> keys;
  std::map<int, Batch> batches;
  char cmd[64];
  int n = 0;
  while (scanf("%63s", cmd) == 1) {
    if (!strcmp(cmd, "joins")) {
      if (scanf("%d", &n) != 1) return 2;
      for (int j = 0; j < n; ++j) {
        long long nb, ns;
        int kb;
        if (scanf("%lld %lld %d", &nb, &ns, &kb) != 3) return 2;
        void* bk = dev_alloc(size_t(nb) * kb, "build keys");
        void* pk = dev_alloc(size_t(ns) * kb, "probe keys");
        keys.push_back(bk);
        keys.push_back(pk);
        tb.push_back(pj_table{bk, nullptr, nb, kb, 0});
        tp.push_back(pj_table{pk, nullptr, ns, kb, 0});
      }
    } else if (!strcmp(cmd, "eager") || !strcmp(cmd, "eager_short")) {
      const bool short_ws = !strcmp(cmd, "eager_short");
      for (int j = 0; j < n; ++j) {
        size_t need = 0;
        expect_ok(pj_workspace_size(&tb[j], &tp[j], nullptr, &need), "pj_workspace_size");
        void* ws = dev_alloc(need < (1u << 20) ? (1u << 20) : need, "eager workspace");
        if (short_ws) {
          fake_unregister(ws);
          fake_register(ws, (need < (1u << 20) ? (1u << 20) : need) - 4096, "eager workspace (short)");
        }
        const int64_t cap = tp[j].n > 0 ? tp[j].n : 1;
        void* out = dev_alloc(size_t(cap) * 8, "eager output");
        int64_t count = 0;
        const pj_status st = pj_partitioned_join(&tb[j], &tp[j], nullptr, static_cast<pj_pair*>(out), cap, &count,
                                                 ws, need < (1u << 20) ? (1u << 20) : need, nullptr);
        // kernels do not run here: the count read back is whatever the buffer held
        if (st != PJ_OK && st != PJ_ERR_CAPACITY) expect_ok(st, "pj_partitioned_join");
        dev_free(out);
        dev_free(ws);
      }
    } else if (!strcmp(cmd, "noalloc")) {
      for (int j = 0; j < n; ++j) {
        const int64_t cap = tp[j].n > 0 ? tp[j].n : 1;
        void* out = dev_alloc(size_t(cap) * 8, "noalloc output");
        void* cnt = dev_alloc(8, "noalloc count");
        const pj_status st = pj_partitioned_join_noalloc(&tb[j], &tp[j], nullptr, static_cast<pj_pair*>(out), cap,
                                                         static_cast<int64_t*>(cnt), nullptr);
        if (st != PJ_ERR_WORKSPACE) expect_ok(st, "pj_partitioned_join_noalloc");
        dev_free(out);
        dev_free(cnt);
      }
    } else if (!strcmp(cmd, "create")) {
      int id, chains;
      char kind[16];
      if (scanf("%d %d %15s", &id, &chains, kind) != 3) return 2;
      Batch B;
      if (!strcmp(kind, "plans")) {
        B.plans.resize(n);
        for (int j = 0; j < n; ++j) {
          pj_plan& p = B.plans[j];
          memset(&p, 0, sizeof(p));
          long long slice;
          if (scanf("%d %d %d %lf %d %lld %d", &p.passes, &p.bits1, &p.bits2, &p.max_partition_size_factor,
                    &p.unique_build_keys, &slice, &p.source) != 7)
            return 2;
          p.probe_slice = slice;
        }
      }
      const pj_plan* plans = B.plans.empty() ? nullptr : B.plans.data();
      int64_t total = 0;
      for (int j = 0; j < n; ++j) {
        B.offs.push_back(total);
        B.caps.push_back(tp[j].n > 0 ? tp[j].n : 1);
        total += B.caps.back();
      }
      size_t need = 0;
      expect_ok(pj_batch_workspace_size(tb.data(), tp.data(), plans, n, chains, &need), "pj_batch_workspace_size");
      B.out = dev_alloc(size_t(total) * 8, "batch output");
      B.counts = dev_alloc(size_t(n) * 8, "batch counts");
      B.ws = dev_alloc(need > 256 ? need : 256, "batch workspace");
      expect_ok(pj_batch_create(tb.data(), tp.data(), plans, n, chains, static_cast<pj_pair*>(B.out), B.offs.data(),
                                B.caps.data(), static_cast<int64_t*>(B.counts), B.ws, need > 256 ? need : 256, &B.h),
                "pj_batch_create");
      batches[id] = std::move(B);
    } else if (!strcmp(cmd, "launch")) {
      int id, reps;
      if (scanf("%d %d", &id, &reps) != 2) return 2;
      for (int r = 0; r < reps; ++r) expect_ok(pj_batch_launch(batches.at(id).h, nullptr), "pj_batch_launch");
    } else if (!strcmp(cmd, "destroy")) {
      int id;
      if (scanf("%d", &id) != 1) return 2;
      Batch& B = batches.at(id);
      expect_ok(pj_batch_destroy(B.h), "pj_batch_destroy");
      dev_free(B.out);
      dev_free(B.counts);
      dev_free(B.ws);
      batches.erase(id);
    } else {
      fprintf(stderr, "[driver] unknown command %s\n", cmd);
      return 2;
    }
  }
  for (auto& kv : batches) {
    pj_batch_destroy(kv.second.h);
    dev_free(kv.second.out);
    dev_free(kv.second.counts);
    dev_free(kv.second.ws);
  }
  for (void* k : keys) dev_free(k);
  printf("launches checked %ld, launch errors %d, call failures %d\n", fake_launches(), fake_errors(), g_fail);
  return (fake_errors() || g_fail) ? 1 : 0;
}
