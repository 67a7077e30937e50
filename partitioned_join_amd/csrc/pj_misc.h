// pj_misc.h -- launchers of pj_misc.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pj.h"

namespace pj {

hipError_t gen_build_unique_i32(int32_t* d, int64_t n, int64_t domain, uint64_t perm_seed, hipStream_t s);
hipError_t gen_build_dup_i32(int32_t* d, int64_t n, int64_t m, int64_t domain, uint64_t perm_seed, uint64_t seed,
                             hipStream_t s);
hipError_t gen_probe_i32(int32_t* d, int64_t n, int64_t domain, int64_t n_build, uint64_t perm_seed, uint64_t seed,
                         double sel, int64_t nm_base, int64_t nm_range, hipStream_t s);
hipError_t gen_probe_zipf_i32(int32_t* d, int64_t count, const double* lut, int64_t n, uint64_t rank_perm_seed,
                              uint64_t seed, hipStream_t s);
hipError_t gen_build_i64(int64_t* d, int64_t n, uint64_t key_seed, hipStream_t s);
hipError_t gen_probe_i64(int64_t* d, int64_t n, int64_t n_build, uint64_t key_seed, uint64_t seed, double sel,
                         hipStream_t s);
hipError_t gen_probe_zipf_i64(int64_t* d, int64_t count, const double* lut, int64_t n_build, uint64_t key_seed,
                              uint64_t rank_perm_seed, uint64_t seed, hipStream_t s);
hipError_t gen_row_ids(int32_t* d, int64_t n, hipStream_t s);

hipError_t hash_keys(const void* keys, int64_t n, int key_bytes, uint32_t* out, hipStream_t s);
hipError_t fastdiv_run(const int32_t* nv, int64_t count, int32_t d, int32_t* q, int32_t* r, hipStream_t s);
hipError_t project_reference(const void* keys, int key_bytes, const pj_pair* pairs, int64_t n, void* out,
                             hipStream_t s);
hipError_t gather_columns(const pj_pair* pairs, int64_t n, const void* bcol, int bbytes, const void* pcol,
                          int pbytes, void* bout, void* pout, hipStream_t s);
hipError_t digit_hist(const void* keys, int64_t n, int key_bytes, int stride, int shift, int bits, uint32_t* hist,
                      hipStream_t s);
size_t scan_ws_bytes(int64_t n);
hipError_t scan_u32(const uint32_t* in, uint32_t* out, int64_t n, int inclusive, uint32_t* ws, hipStream_t s);
hipError_t assign_blocks(const uint32_t* off, int64_t total, int P, int base, uint32_t* scanned, uint32_t* pidx,
                         int64_t max_blocks, uint32_t* d_total, uint32_t* ws, hipStream_t s);
hipError_t p2_prep(const uint32_t* off1, int F1, uint32_t nseg, uint32_t* matc, uint32_t* mat_row0, int P,
                   hipStream_t s);

}  // namespace pj
