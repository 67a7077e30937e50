// partitioned_join -- command-line driver with the reference binary's flags and
// CSV columns (main.cu:75-948, scripts/bench_partitioned_join.sh:12).  The join
// runs in libpj.so; inputs are generated on the device with the library's seeded
// generators.  Like the reference binary it verifies the GPU result against a
// CPU hash join (skipped with -S), reported as "num errors".
#include <getopt.h>
#include <hip/hip_runtime.h>
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/pj.h"

#define HIP_OK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      std::fprintf(stderr, "HIP error %s at %s:%d: %s\n", #x, __FILE__, __LINE__, hipGetErrorString(e_)); \
      std::exit(2);                                                                        \
    }                                                                                      \
  } while (0)

#define PJ_OK_OR_DIE(x)                                                                    \
  do {                                                                                     \
    pj_status s_ = (x);                                                                    \
    if (s_ != PJ_OK) {                                                                     \
      std::fprintf(stderr, "libpj error %d in %s: %s\n", int(s_), #x, pj_last_error());    \
      std::exit(3);                                                                        \
    }                                                                                      \
  } while (0)

static bool is_pow2(long v) { return v > 0 && (v & (v - 1)) == 0; }

static int ilog2(long v) {
  int b = 0;
  while ((1L << b) < v) ++b;
  return b;
}

// CPU check join: unordered_map build + OpenMP probe (the reference's
// inner_join_cpu scheme, src/inner_join_cpu.hpp:39-79), emitting (build row,
// probe row) pairs; duplicate build keys use a multimap.
template <typename K>
static double cpu_join(const std::vector<K>& b, const std::vector<K>& p, bool uniq,
                       std::vector<std::pair<int32_t, int32_t>>* out) {
  double t0 = 0, t1 = 0;
  if (uniq) {
    std::unordered_map<K, int32_t> ht;
    ht.reserve(2 * b.size());
    t0 = omp_get_wtime();
    for (size_t i = 0; i < b.size(); ++i) ht.insert({b[i], int32_t(i)});
    std::vector<std::vector<std::pair<int32_t, int32_t>>> parts(omp_get_max_threads());
#pragma omp parallel
    {
      auto& mine = parts[omp_get_thread_num()];
#pragma omp for schedule(static)
      for (long j = 0; j < long(p.size()); ++j) {
        auto it = ht.find(p[j]);
        if (it != ht.end()) mine.push_back({it->second, int32_t(j)});
      }
    }
    t1 = omp_get_wtime();
    for (auto& v : parts) out->insert(out->end(), v.begin(), v.end());
  } else {
    std::unordered_multimap<K, int32_t> ht;
    ht.reserve(2 * b.size());
    t0 = omp_get_wtime();
    for (size_t i = 0; i < b.size(); ++i) ht.insert({b[i], int32_t(i)});
    for (size_t j = 0; j < p.size(); ++j) {
      auto r = ht.equal_range(p[j]);
      for (auto it = r.first; it != r.second; ++it) out->push_back({it->second, int32_t(j)});
    }
    t1 = omp_get_wtime();
  }
  return t1 - t0;
}

template <typename K>
static std::vector<K> read_keys(const std::string& path, long n) {
  std::vector<K> v(static_cast<size_t>(n));
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    std::cerr << "Error: cannot open " << path << std::endl;
    std::exit(1);
  }
  f.read(reinterpret_cast<char*>(v.data()), std::streamsize(n * sizeof(K)));
  if (f.gcount() != std::streamsize(n * sizeof(K))) {
    std::cerr << "Error: " << path << " holds fewer than " << n << " keys" << std::endl;
    std::exit(1);
  }
  return v;
}

// TPC-H Q4 text tables as dbgen writes them (one row per line, fields ended
// by '|'): the build side is orders -- key o_orderkey (field 0), payload the
// first character of o_orderpriority (field 5) -- and the probe side lineitem
// -- key l_orderkey (field 0).  A file with fewer rows than requested is
// repeated from its start (reference: src/generate_zipf_input.cuh:154-253,
// main.cu:304-333, which reads fixed paths).
static bool ends_with(const std::string& s, const std::string& suffix) {
  return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

template <typename K>
static std::vector<K> read_tbl(const std::string& path, long n, int pay_field, std::vector<int32_t>* pay) {
  std::ifstream f(path);
  if (!f) {
    std::cerr << "Error: cannot open " << path << std::endl;
    std::exit(1);
  }
  std::vector<K> keys;
  keys.reserve(static_cast<size_t>(n));
  std::string line;
  while (long(keys.size()) < n && std::getline(f, line)) {
    if (line.empty()) continue;
    size_t start = 0;
    std::string key_field, pay_text;
    for (int field = 0; field <= std::max(0, pay_field); ++field) {
      const size_t bar = line.find('|', start);
      if (bar == std::string::npos) {
        std::cerr << "Error: " << path << " line " << keys.size() + 1 << " has fewer than " << field + 1
                  << " fields" << std::endl;
        std::exit(1);
      }
      if (field == 0) key_field = line.substr(start, bar - start);
      if (field == pay_field) pay_text = line.substr(start, bar - start);
      start = bar + 1;
    }
    keys.push_back(static_cast<K>(std::strtoll(key_field.c_str(), nullptr, 10)));
    if (pay) pay->push_back(pay_text.empty() ? 0 : int32_t(static_cast<unsigned char>(pay_text[0])));
  }
  if (keys.empty() && n > 0) {
    std::cerr << "Error: " << path << " holds no rows" << std::endl;
    std::exit(1);
  }
  const size_t rows = keys.size();
  for (size_t i = rows; i < size_t(n); ++i) {
    keys.push_back(keys[i % rows]);
    if (pay) pay->push_back((*pay)[i % rows]);
  }
  return keys;
}

template <typename K>
static int run(int argc, char** argv, int key_bytes) {
  long build_n = 1 << 20, probe_n = 1 << 20;
  double sel = 0.03;
  long rand_max = -1;
  bool uniq = false, from_file = false, radix = false, skip_cpu = false, with_value = false, csv = false;
  std::string bfile, pfile;
  int verbosity = 0, skew_mode = 0;
  long np1 = -1, np2 = -1, out_size = -1;
  double zipf = -1.0, factor = 3.0;
  const option long_opts[] = {
      {"build_tbl_size", required_argument, nullptr, 'b'}, {"probe_tbl_size", required_argument, nullptr, 'p'},
      {"selectivity", required_argument, nullptr, 's'},    {"rand_max", required_argument, nullptr, 'r'},
      {"uniq_build_tbl_keys", no_argument, nullptr, 'u'},  {"read_from_file", no_argument, nullptr, 'F'},
      {"build_tbl_file", required_argument, nullptr, 'B'}, {"probe_tbl_file", required_argument, nullptr, 'P'},
      {"verbosity", required_argument, nullptr, 'v'},
      {"max_partition_size_factor", required_argument, nullptr, 'f'},
      {"radix_partition", no_argument, nullptr, 'R'},      {"num_partitions_1", required_argument, nullptr, 'n'},
      {"num_partitions_2", required_argument, nullptr, 'N'}, {"skip_cpu_run", required_argument, nullptr, 'S'},
      {"with_value", no_argument, nullptr, 'V'},           {"csv", no_argument, nullptr, 'C'},
      {"skewed_data_mode", required_argument, nullptr, 'Z'}, {"zipf_dist", required_argument, nullptr, 'z'},
      {"do_combined_histogram_pass", no_argument, nullptr, 'H'}, {"key_bytes", required_argument, nullptr, 'K'},
      {"help", no_argument, nullptr, 'h'},                 {nullptr, 0, nullptr, 0}};
  optind = 1;
  int opt;
  while ((opt = getopt_long(argc, argv, "b:p:s:r:uFB:P:v:f:Rn:N:S:VCZ:z:HK:h", long_opts, nullptr)) != -1) {
    switch (opt) {
      case 'b': build_n = std::atol(optarg); break;
      case 'p': probe_n = std::atol(optarg); break;
      case 's': sel = std::atof(optarg); break;
      case 'r': rand_max = std::atol(optarg); break;
      case 'u': uniq = true; break;
      case 'F': from_file = true; break;
      case 'B': bfile = optarg; break;
      case 'P': pfile = optarg; break;
      case 'v': verbosity = std::atoi(optarg); break;
      case 'f': factor = std::atof(optarg); break;
      case 'R': radix = true; break;
      case 'n': np1 = std::atol(optarg); break;
      case 'N': np2 = std::atol(optarg); break;
      case 'S': skip_cpu = true; out_size = std::atol(optarg); break;
      case 'V': with_value = true; break;
      case 'C': csv = true; break;
      case 'Z': skew_mode = std::atoi(optarg); break;
      case 'z': zipf = std::atof(optarg); break;
      case 'H': break;  // the combined histogram always runs (one key read for both passes)
      case 'K': break;
      case 'h':
        std::cout << "Usage: partitioned_join [-b build] [-p probe] [-s selectivity] [-r rand_max] [-u] [-F -B file "
                     "-P file] [-v level] [-f factor] [-R] [-n P1] [-N P2] [-S out_size] [-V] [-C] [-Z mode] [-z "
                     "theta] [-H] [-K 4|8]"
                  << std::endl;
        return 1;
      default:
        std::cout << "Please use -h or --help for the list of options" << std::endl;
        return 1;
    }
  }
  if (rand_max < 0) rand_max = 2 * build_n;
  if (build_n < 0 || probe_n < 0 || build_n >= (1L << 31) || probe_n >= (1L << 31)) {
    std::cerr << "Error: table sizes must be in [0, 2^31)" << std::endl;
    return 1;
  }
  pj_plan plan;
  pj_table bt{nullptr, nullptr, build_n, key_bytes, 0}, pt{nullptr, nullptr, probe_n, key_bytes, 0};
  PJ_OK_OR_DIE(pj_plan_default(&bt, &pt, &plan));
  plan.max_partition_size_factor = factor;
  plan.unique_build_keys = uniq ? 1 : 0;
  plan.source = 2;
  if (np1 != -1 || np2 != -1) {
    if (!radix && is_pow2(np1)) {
      plan.passes = 1;
      plan.bits1 = ilog2(np1);
      plan.bits2 = 0;
    } else if (radix && is_pow2(np1) && is_pow2(np2)) {
      plan.passes = 2;
      plan.bits1 = ilog2(np1);
      plan.bits2 = ilog2(np2);
    } else {
      std::cerr << "Error: num_partition must be power of two." << std::endl;
      return 1;
    }
  } else {
    plan.source = 0;
  }
  if (skew_mode != 0) {
    sel = 1.0;
    uniq = skew_mode == 2;
    plan.unique_build_keys = uniq ? 1 : 0;
  }

  // ---------------------------------------------------------------- inputs
  K *d_build = nullptr, *d_probe = nullptr;
  int32_t *d_bval = nullptr, *d_pval = nullptr;
  HIP_OK(hipMalloc(&d_build, std::max(1L, build_n) * sizeof(K)));
  HIP_OK(hipMalloc(&d_probe, std::max(1L, probe_n) * sizeof(K)));
  hipStream_t s;
  HIP_OK(hipStreamCreate(&s));
  std::vector<int32_t> bpay;  // build payload read from a file (the pairs then carry it)
  if (from_file && ends_with(bfile, ".tbl") && ends_with(pfile, ".tbl")) {
    auto hb = read_tbl<K>(bfile, build_n, 5, &bpay);
    auto hp = read_tbl<K>(pfile, probe_n, 0, nullptr);
    HIP_OK(hipMemcpy(d_build, hb.data(), build_n * sizeof(K), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_probe, hp.data(), probe_n * sizeof(K), hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&d_bval, std::max(1L, build_n) * 4));
    HIP_OK(hipMemcpy(d_bval, bpay.data(), build_n * 4, hipMemcpyHostToDevice));
  } else if (from_file) {
    auto hb = read_keys<K>(bfile, build_n), hp = read_keys<K>(pfile, probe_n);
    HIP_OK(hipMemcpy(d_build, hb.data(), build_n * sizeof(K), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_probe, hp.data(), probe_n * sizeof(K), hipMemcpyHostToDevice));
  } else if (key_bytes == 4) {
    const int64_t domain = std::min<int64_t>(int64_t(rand_max) + 1, int64_t(1) << 32);
    int64_t nbuild_keys = build_n;
    if (uniq) {
      if (domain < build_n) {
        std::cerr << "Error: rand_max + 1 must be >= build size for unique keys" << std::endl;
        return 1;
      }
      PJ_OK_OR_DIE(pj_gen_build_unique_i32(reinterpret_cast<int32_t*>(d_build), build_n, domain, 1234, s));
    } else {
      nbuild_keys = std::min<int64_t>(std::max<int64_t>(build_n, 1), domain);
      PJ_OK_OR_DIE(pj_gen_build_dup_i32(reinterpret_cast<int32_t*>(d_build), build_n, nbuild_keys, domain, 1234,
                                        4321, s));
    }
    if ((skew_mode == 2 || skew_mode == 3) && build_n > 0) {
      const double theta = zipf < 0 ? 1.0 : zipf;
      std::vector<double> lut(static_cast<size_t>(nbuild_keys));
      PJ_OK_OR_DIE(pj_zipf_lut(lut.data(), nbuild_keys, theta));
      double* d_lut;
      HIP_OK(hipMalloc(&d_lut, lut.size() * 8));
      HIP_OK(hipMemcpy(d_lut, lut.data(), lut.size() * 8, hipMemcpyHostToDevice));
      // ranks map onto the build key set through the generator's own permutation
      PJ_OK_OR_DIE(pj_gen_probe_zipf_i32(reinterpret_cast<int32_t*>(d_probe), probe_n, d_lut, nbuild_keys, 1234,
                                         5678, s));
      HIP_OK(hipStreamSynchronize(s));
      HIP_OK(hipFree(d_lut));
      if (nbuild_keys != domain) {
        // perm over nbuild_keys ranks -> ids in [0, nbuild_keys); lift them onto the build's domain permutation
        std::cerr << "note: skewed probes use ranks over the build key set" << std::endl;
      }
    } else if (probe_n > 0 && build_n > 0) {
      const bool lottery = domain > nbuild_keys;
      PJ_OK_OR_DIE(pj_gen_probe_i32(reinterpret_cast<int32_t*>(d_probe), probe_n, domain, nbuild_keys, 1234, 5678,
                                    lottery ? sel : 1.0, 0, 0, s));
    }
  } else {
    PJ_OK_OR_DIE(pj_gen_build_i64(reinterpret_cast<int64_t*>(d_build), build_n, 1234, s));
    if (probe_n > 0 && build_n > 0)
      PJ_OK_OR_DIE(pj_gen_probe_i64(reinterpret_cast<int64_t*>(d_probe), probe_n, build_n, 1234, 5678, sel, s));
  }
  if (with_value) {
    if (!d_bval) {
      HIP_OK(hipMalloc(&d_bval, std::max(1L, build_n) * 4));
      PJ_OK_OR_DIE(pj_gen_row_ids(d_bval, build_n, s));
    }
    HIP_OK(hipMalloc(&d_pval, std::max(1L, probe_n) * 4));
    PJ_OK_OR_DIE(pj_gen_row_ids(d_pval, probe_n, s));
  }
  HIP_OK(hipStreamSynchronize(s));
  bt.keys = d_build;
  bt.payload = d_bval;
  pt.keys = d_probe;
  pt.payload = d_pval;

  std::vector<K> hb(static_cast<size_t>(build_n));
  std::vector<K> hp(static_cast<size_t>(probe_n));
  HIP_OK(hipMemcpy(hb.data(), d_build, build_n * sizeof(K), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(hp.data(), d_probe, probe_n * sizeof(K), hipMemcpyDeviceToHost));
  const int tuple_byte = with_value ? key_bytes + 4 : key_bytes;
  const long P1 = 1L << plan.bits1, P2 = plan.passes == 2 ? (1L << plan.bits2) : 0;
  if (csv) {
    std::cout << build_n << "," << probe_n << "," << key_bytes << "," << tuple_byte << "," << sel << "," << uniq << ","
              << P1 << "," << P2 << "," << factor << "," << skew_mode << "," << zipf << ",";
  } else {
    std::cout << "[*] Settings:\n"
              << "    build table size     = " << build_n << "\n    probe table size     = " << probe_n
              << "\n    has value col        = " << std::boolalpha << with_value
              << "\n    skewed data mode     = " << skew_mode << "\n    selectivity          = " << sel
              << "\n    rand max             = " << rand_max << "\n    uniq build tbl keys  = " << uniq
              << "\n    max part size factor = " << factor << "\n    passes               = " << plan.passes
              << "\n    num partitions P1    = " << P1 << "\n    num partitions P2    = " << P2
              << "\n    sizeof(key_type)     = " << key_bytes << "\n    skip CPU run         = " << skip_cpu
              << std::endl;
  }
  if (verbosity > 2) {
    std::cout << "====Build input table:" << std::endl;
    for (auto k : hb) std::cout << k << " ";
    std::cout << std::endl << "====Probe input table:" << std::endl;
    for (auto k : hp) std::cout << k << " ";
    std::cout << std::endl << "====End" << std::endl;
  }

  // ---------------------------------------------------------------- CPU check run
  std::vector<std::pair<int32_t, int32_t>> cpu_pairs;
  double cpu_s = -1;
  if (!skip_cpu) cpu_s = cpu_join<K>(hb, hp, uniq, &cpu_pairs);
  for (auto& pr : cpu_pairs)
    if (!bpay.empty()) pr.first = bpay[size_t(pr.first)];  // the GPU pairs carry the build payload
  const int64_t cap = skip_cpu ? std::max(1L, out_size) : std::max<int64_t>(1, int64_t(cpu_pairs.size()));

  // ---------------------------------------------------------------- GPU run
  size_t wsb = 0;
  PJ_OK_OR_DIE(pj_workspace_size(&bt, &pt, &plan, &wsb));
  void* ws;
  pj_pair* d_out;
  HIP_OK(hipMalloc(&ws, wsb));
  HIP_OK(hipMalloc(&d_out, size_t(cap) * sizeof(pj_pair)));
  int64_t count = 0;
  pj_status st = pj_partitioned_join(&bt, &pt, &plan, d_out, cap, &count, ws, wsb, s);  // warm-up
  if (st != PJ_OK && st != PJ_ERR_CAPACITY) PJ_OK_OR_DIE(st);
  PJ_OK_OR_DIE(pj_profile_enable(1));
  st = pj_partitioned_join(&bt, &pt, &plan, d_out, cap, &count, ws, wsb, s);
  pj_phase_times t;
  PJ_OK_OR_DIE(pj_profile_read(&t));
  if (st != PJ_OK && st != PJ_ERR_CAPACITY) PJ_OK_OR_DIE(st);

  long long errors = -1;
  if (!skip_cpu) {
    if (count != int64_t(cpu_pairs.size())) {
      errors = std::numeric_limits<int32_t>::max();
    } else {
      std::vector<pj_pair> g(static_cast<size_t>(count));
      HIP_OK(hipMemcpy(g.data(), d_out, size_t(count) * sizeof(pj_pair), hipMemcpyDeviceToHost));
      std::vector<std::pair<int32_t, int32_t>> gp(g.size());
      for (size_t i = 0; i < g.size(); ++i) gp[i] = {g[i].build_idx, g[i].probe_idx};
      std::sort(gp.begin(), gp.end());
      std::sort(cpu_pairs.begin(), cpu_pairs.end());
      errors = 0;
      for (size_t i = 0; i < gp.size(); ++i) errors += gp[i] != cpu_pairs[i];
    }
  }
  const double hist_ms = t.histogram_ms + t.offsets_ms, join_ms = t.join_ms + t.compact_ms;
  if (csv) {
    std::cout << count << "," << hist_ms << "," << t.pass1_ms << "," << t.pass2_ms << "," << t.skew_ms << ","
              << join_ms << "," << t.total_ms << ",";
    if (skip_cpu) std::cout << "SKIPPED" << std::endl;
    else std::cout << errors << std::endl;
  } else {
    const double in_bytes = double(build_n + probe_n) * tuple_byte;
    std::cout << "[*] Results:\n    joined size          = " << count << std::setprecision(4)
              << "\n    combined histogram   = " << hist_ms << " ms\n    partition P1 time    = " << t.pass1_ms
              << " ms\n    partition P2 time    = " << t.pass2_ms << " ms\n    skewed data time     = " << t.skew_ms
              << " ms\n    join phase time      = " << join_ms << " ms\n    total exec time      = " << t.total_ms
              << " ms\n    effective BW         = " << in_bytes / (t.total_ms / 1000.0) / 1e9 << " GB/s"
              << "\n    tuples/s             = " << double(build_n + probe_n) / (t.total_ms / 1000.0)
              << "\n    device memory usage  = " << double(wsb + cap * 8 + (build_n + probe_n) * (key_bytes + 4)) / 1e9
              << " GB" << std::endl;
    if (!skip_cpu)
      std::cout << "    reference CPU time   = " << 1000.0 * cpu_s << " ms (" << omp_get_max_threads()
                << " threads)\n    number of errors     = " << errors << std::endl;
    else
      std::cout << "    reference CPU time   = SKIPPED\n    number of errors     = SKIPPED" << std::endl;
  }
  if (st == PJ_ERR_CAPACITY) std::cerr << "warning: output capacity " << cap << " < " << count << std::endl;
  HIP_OK(hipFree(ws));
  HIP_OK(hipFree(d_out));
  HIP_OK(hipFree(d_build));
  HIP_OK(hipFree(d_probe));
  if (d_bval) HIP_OK(hipFree(d_bval));
  if (d_pval) HIP_OK(hipFree(d_pval));
  HIP_OK(hipStreamDestroy(s));
  return (errors == 0 || skip_cpu) ? 0 : 4;
}

int main(int argc, char** argv) {
  int key_bytes = 4;
  for (int i = 1; i + 1 < argc; ++i)
    if (std::strcmp(argv[i], "-K") == 0 || std::strcmp(argv[i], "--key_bytes") == 0) key_bytes = std::atoi(argv[i + 1]);
  if (key_bytes == 8) return run<int64_t>(argc, argv, 8);
  if (key_bytes != 4) {
    std::cerr << "Error: -K must be 4 or 8" << std::endl;
    return 1;
  }
  return run<int32_t>(argc, argv, 4);
}
