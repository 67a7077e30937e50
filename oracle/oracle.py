"""CPU oracle for the radix-partitioned join path.

TEST INFRASTRUCTURE ONLY: the checker.  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg may import this module.  The product
(``partitioned_join_amd`` / ``libpj.so``) never imports, links or calls it.

Two layers:

* ``liboracle.so`` (``oracle/pj_oracle.cpp``): C++ restatement of the reference's
  CPU join (``src/inner_join_cpu.hpp:23-107``), MurmurHash3
  (``src/hash_functions.cuh:33-125``), int_fastdiv (``src/int_fastdiv.h:21-133``),
  the seeded generators, a sort-merge join and per-phase restatements.
* numpy restatements of the phase arithmetic below (histogram digits, scans,
  block assignment, scatter checks), each citing the reference line it follows.

The reference's own CPU code is not built here: its headers need the CUDA
toolkit (``cuda_runtime.h``, ``int2`` and Thrust), which this image lacks, so the
restatement is pinned by mathematics instead (SMHasher's verification value, C
division semantics, brute-force and sort-merge joins, closed forms and cuDF's
published inner-join answers; see ``tests/test_oracle.py``).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

c_i32, c_i64, c_u32, c_u64, c_int, c_dbl, c_vp = (
    ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
    ctypes.c_double, ctypes.c_void_p)


class FastDiv(ctypes.Structure):
    _fields_ = [("d", c_i32), ("magic", c_i32), ("shift", c_i32), ("add_sign", c_i32)]


def lib() -> ctypes.CDLL:
    """Load oracle/liboracle.so (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        sig = {
            "pjo_murmur3_32": (c_u32, [c_vp, c_int, c_u32]),
            "pjo_murmur3_key4": (c_u32, [c_u32]),
            "pjo_murmur3_key8": (c_u32, [c_u64]),
            "pjo_fastdiv_init": (None, [ctypes.POINTER(FastDiv), c_i32]),
            "pjo_fastdiv_div": (c_i32, [c_i32, ctypes.POINTER(FastDiv)]),
            "pjo_fastdiv_mod": (c_i32, [c_i32, ctypes.POINTER(FastDiv)]),
            "pjo_hash_slot": (c_i32, [c_u32, c_i32]),
            "pjo_mix64": (c_u64, [c_u64]),
            "pjo_rand64": (c_u64, [c_u64, c_u64]),
            "pjo_bounded": (c_u32, [c_u64, c_u64]),
            "pjo_perm": (c_u64, [c_u64, c_u64, c_u64]),
            "pjo_gen_build_unique_i32": (None, [c_vp, c_i64, c_i64, c_u64]),
            "pjo_gen_build_dup_i32": (None, [c_vp, c_i64, c_i64, c_i64, c_u64, c_u64]),
            "pjo_gen_probe_i32": (None, [c_vp, c_i64, c_i64, c_i64, c_u64, c_u64, c_dbl, c_i64, c_i64]),
            "pjo_zipf_lut": (None, [c_vp, c_i64, c_dbl]),
            "pjo_zipf_rank": (c_i64, [c_vp, c_i64, c_dbl]),
            "pjo_gen_probe_zipf_i32": (None, [c_vp, c_i64, c_vp, c_i64, c_u64, c_u64]),
            "pjo_gen_build_i64": (None, [c_vp, c_i64, c_u64]),
            "pjo_gen_probe_i64": (None, [c_vp, c_i64, c_i64, c_u64, c_u64, c_dbl]),
            "pjo_gen_probe_zipf_i64": (None, [c_vp, c_i64, c_vp, c_i64, c_u64, c_u64, c_u64]),
            "pjo_inner_join_cpu": (c_i64, [c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, c_int, c_int, c_int,
                                           c_vp, c_i64, ctypes.POINTER(c_dbl)]),
            "pjo_sort_merge_join": (c_i64, [c_vp, c_vp, c_i64, c_vp, c_vp, c_i64, c_int, c_vp, c_i64]),
            "pjo_join_count": (c_i64, [c_vp, c_i64, c_vp, c_i64, c_int]),
            "pjo_sort_pairs": (None, [c_vp, c_i64]),
            "pjo_histogram": (None, [c_vp, c_i64, c_int, c_int, c_int, c_vp]),
            "pjo_partition_pass": (None, [c_vp, c_vp, c_i64, c_int, c_int, c_int, c_vp, c_vp, c_vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(c_vp)


# ---------------------------------------------------------------- hashing
def murmur3_32(data: bytes, seed: int = 0) -> int:
    buf = ctypes.create_string_buffer(bytes(data), len(data))
    return lib().pjo_murmur3_32(ctypes.cast(buf, c_vp), len(data), seed)


def murmur3_key(key: int, key_bytes: int = 4) -> int:
    if key_bytes == 8:
        return lib().pjo_murmur3_key8(key & 0xFFFFFFFFFFFFFFFF)
    return lib().pjo_murmur3_key4(key & 0xFFFFFFFF)


def fastdiv(d: int) -> FastDiv:
    f = FastDiv()
    lib().pjo_fastdiv_init(ctypes.byref(f), d)
    return f


def fastdiv_div(n: int, f: FastDiv) -> int:
    return lib().pjo_fastdiv_div(n, ctypes.byref(f))


def fastdiv_mod(n: int, f: FastDiv) -> int:
    return lib().pjo_fastdiv_mod(n, ctypes.byref(f))


def hash_slot(h: int, table_size: int) -> int:
    return lib().pjo_hash_slot(h, table_size)


# ---------------------------------------------------------------- generators
def gen_build_unique_i32(count: int, domain: int, perm_seed: int) -> np.ndarray:
    out = np.empty(count, dtype=np.int32)
    lib().pjo_gen_build_unique_i32(_ptr(out), count, domain, perm_seed)
    return out


def gen_build_dup_i32(count: int, m: int, domain: int, perm_seed: int, seed: int) -> np.ndarray:
    out = np.empty(count, dtype=np.int32)
    lib().pjo_gen_build_dup_i32(_ptr(out), count, m, domain, perm_seed, seed)
    return out


def gen_probe_i32(count: int, domain: int, n_build: int, perm_seed: int, seed: int, sel: float,
                  nonmatch_base: int = 0, nonmatch_range: int = 0) -> np.ndarray:
    out = np.empty(count, dtype=np.int32)
    lib().pjo_gen_probe_i32(_ptr(out), count, domain, n_build, perm_seed, seed, sel,
                            nonmatch_base, nonmatch_range)
    return out


def zipf_lut(n: int, theta: float) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().pjo_zipf_lut(_ptr(out), n, theta)
    return out


def zipf_rank(lut: np.ndarray, u: float) -> int:
    return lib().pjo_zipf_rank(_ptr(lut), len(lut), u)


def gen_probe_zipf_i32(count: int, lut: np.ndarray, rank_perm_seed: int, seed: int) -> np.ndarray:
    out = np.empty(count, dtype=np.int32)
    lib().pjo_gen_probe_zipf_i32(_ptr(out), count, _ptr(lut), len(lut), rank_perm_seed, seed)
    return out


def gen_build_i64(count: int, key_seed: int) -> np.ndarray:
    out = np.empty(count, dtype=np.int64)
    lib().pjo_gen_build_i64(_ptr(out), count, key_seed)
    return out


def gen_probe_i64(count: int, n_build: int, key_seed: int, seed: int, sel: float) -> np.ndarray:
    out = np.empty(count, dtype=np.int64)
    lib().pjo_gen_probe_i64(_ptr(out), count, n_build, key_seed, seed, sel)
    return out


def gen_probe_zipf_i64(count: int, lut: np.ndarray, key_seed: int, rank_perm_seed: int,
                       seed: int) -> np.ndarray:
    out = np.empty(count, dtype=np.int64)
    lib().pjo_gen_probe_zipf_i64(_ptr(out), count, _ptr(lut), len(lut), key_seed, rank_perm_seed, seed)
    return out


# ---------------------------------------------------------------- joins
def _keys_and_width(keys: np.ndarray):
    keys = np.ascontiguousarray(keys)
    if keys.dtype.itemsize == 8:
        return keys.astype(np.int64, copy=False), 8
    return keys.astype(np.int32, copy=False), 4


def _pay(p):
    return None if p is None else np.ascontiguousarray(p, dtype=np.int32)


def inner_join_cpu(build_keys, probe_keys, build_pay=None, probe_pay=None, uniq_keys=True,
                   num_threads=0, capacity=None, return_seconds=False):
    """Restatement of inner_join_cpu (src/inner_join_cpu.hpp:23-107) emitting
    (build payload, probe payload) pairs as an (n, 2) int32 array."""
    bk, kb = _keys_and_width(build_keys)
    pk, _ = _keys_and_width(probe_keys)
    bp, pp = _pay(build_pay), _pay(probe_pay)
    if capacity is None:
        capacity = join_count(bk, pk)
    out = np.empty((max(capacity, 1), 2), dtype=np.int32)
    secs = c_dbl(0.0)
    n = lib().pjo_inner_join_cpu(_ptr(bk), _ptr(bp), len(bk), _ptr(pk), _ptr(pp), len(pk), kb,
                                 int(bool(uniq_keys)), num_threads, _ptr(out), capacity, ctypes.byref(secs))
    res = out[: min(n, capacity)]
    return (res, n, secs.value) if return_seconds else res


def sort_merge_join(build_keys, probe_keys, build_pay=None, probe_pay=None):
    bk, kb = _keys_and_width(build_keys)
    pk, _ = _keys_and_width(probe_keys)
    bp, pp = _pay(build_pay), _pay(probe_pay)
    cap = join_count(bk, pk)
    out = np.empty((max(cap, 1), 2), dtype=np.int32)
    n = lib().pjo_sort_merge_join(_ptr(bk), _ptr(bp), len(bk), _ptr(pk), _ptr(pp), len(pk), kb, _ptr(out), cap)
    assert n == cap
    return out[:n]


def join_count(build_keys, probe_keys) -> int:
    bk, kb = _keys_and_width(build_keys)
    pk, _ = _keys_and_width(probe_keys)
    return int(lib().pjo_join_count(_ptr(bk), len(bk), _ptr(pk), len(pk), kb))


def sort_pairs(pairs: np.ndarray) -> np.ndarray:
    """Sorted copy of an (n, 2) int32 pair array, ordered by (first, second)."""
    p = np.ascontiguousarray(pairs, dtype=np.int32).copy()
    if len(p):
        lib().pjo_sort_pairs(_ptr(p), len(p))
    return p


def nested_loop_join(build_keys, probe_keys, build_pay=None, probe_pay=None) -> np.ndarray:
    """Brute force {(b, p) : R.key[b] == S.key[p]} for tiny inputs (textbook definition)."""
    out = []
    for b, kb in enumerate(np.asarray(build_keys).tolist()):
        for p, kp in enumerate(np.asarray(probe_keys).tolist()):
            if kb == kp:
                out.append((b if build_pay is None else int(build_pay[b]),
                            p if probe_pay is None else int(probe_pay[p])))
    return np.array(out, dtype=np.int32).reshape(-1, 2)


def project_to_reference(pairs: np.ndarray, build_keys, probe_pay=None) -> np.ndarray:
    """pi(b, p) = (R.key[b], S.payload[p]): the reference's output form
    {x = key, y = probe payload} (src/hash_join_kernels.cuh:155-164).  ``pairs``
    must hold row ids in column 0 (build) and probe payloads in column 1."""
    keys = np.asarray(build_keys)[pairs[:, 0]]
    y = pairs[:, 1] if probe_pay is None else np.asarray(probe_pay)[pairs[:, 1]]
    return np.stack([keys.astype(np.int64), y.astype(np.int64)], axis=1)


# ---------------------------------------------------------------- phases (numpy)
def radix_digit(keys: np.ndarray, shift: int, bits: int) -> np.ndarray:
    """Digit of each key: pass 1 uses ((P1-1) << b2 & key) >> b2
    (src/partition_kernels.cuh:168,177,236,244,292); pass 2 and the combined
    histogram use key & (P-1) (:109,380,389,471,479,527).  Both are
    (uint(key) >> shift) & (2^bits - 1) over the key's low 32 bits."""
    k = np.asarray(keys).astype(np.int64) & 0xFFFFFFFF
    return ((k >> shift) & ((1 << bits) - 1)).astype(np.int64)


def histogram(keys: np.ndarray, shift: int, bits: int) -> np.ndarray:
    """Closed form bincount of the digit (K0b, K1, K5: src/partition_kernels.cuh:85-124,140-194,328-406)."""
    return np.bincount(radix_digit(keys, shift, bits), minlength=1 << bits).astype(np.int64)


def fold_histogram(fine: np.ndarray, p2: int) -> np.ndarray:
    """retrieve_histogram_of_pass_one: hist1[i / P2] += fine[i] (src/partition_kernels.cuh:126-138)."""
    return fine.reshape(-1, p2).sum(axis=1)


def exclusive_scan(x: np.ndarray) -> np.ndarray:
    """cub::DeviceScan::ExclusiveSum (src/partitioned_join.cuh:86,152)."""
    x = np.asarray(x, dtype=np.int64)
    out = np.zeros_like(x)
    if len(x) > 1:
        out[1:] = np.cumsum(x[:-1])
    return out


def inclusive_scan(x: np.ndarray) -> np.ndarray:
    """cub::DeviceScan::InclusiveSum (src/skewed_data_handler.cuh:79-81,90)."""
    return np.cumsum(np.asarray(x, dtype=np.int64))


def blocks_per_partition(offsets: np.ndarray, total: int, base: int) -> np.ndarray:
    """get_number_of_blocks_each_partition: max(1, ceil(size/base)) with sizes
    from scanned offsets (src/skewed_data_handler.cuh:20-43)."""
    offsets = np.asarray(offsets, dtype=np.int64)
    ends = np.append(offsets[1:], total)
    size = ends - offsets
    return np.where(size == 0, 1, (size - 1) // base + 1).astype(np.int64)


def assign_blocks(offsets: np.ndarray, total: int, base: int):
    """handle_skewed_data (src/skewed_data_handler.cuh:58-91): returns
    (blocks_scanned inclusive, partition_index[block], total_blocks).
    partition_index = repeat(arange(P), blocks) (populate_index_array + InclusiveSum)."""
    nb = blocks_per_partition(offsets, total, base)
    scanned = inclusive_scan(nb)
    part_index = np.repeat(np.arange(len(nb), dtype=np.int64), nb)
    return scanned, part_index, int(scanned[-1]) if len(scanned) else 0


def block_chunk(offsets: np.ndarray, total: int, base: int, scanned: np.ndarray, part_index: np.ndarray,
                block: int):
    """Chunk [start + off*base, min(start + (off+1)*base, end)) handled by a block
    (src/hash_join_kernels.cuh:56-71; src/partition_kernels.cuh:351-363)."""
    p = int(part_index[block])
    start = int(offsets[p])
    end = int(offsets[p + 1]) if p + 1 < len(offsets) else total
    prev = 0 if p == 0 else int(scanned[p - 1])
    off = block - prev
    s = start + off * base
    return s, min(s + base, end)


def check_partitioned(keys_in, pay_in, keys_out, pay_out, offsets, shift, bits) -> None:
    """Scatter pins: (i) output is a permutation of the input tuples, (ii) every
    tuple in [off[p], off[p+1]) has digit p, (iii) off = exclusive scan of the
    histogram.  Intra-partition order is unspecified and not checked."""
    keys_in = np.asarray(keys_in)
    keys_out = np.asarray(keys_out)
    n = len(keys_in)
    pay_in = np.arange(n, dtype=np.int32) if pay_in is None else np.asarray(pay_in)
    hist = histogram(keys_in, shift, bits)
    off = np.asarray(offsets, dtype=np.int64)[: 1 << bits]
    assert np.array_equal(off, exclusive_scan(hist)), "offsets != exclusive scan of histogram"
    d = radix_digit(keys_out, shift, bits)
    expect = np.repeat(np.arange(1 << bits), hist)
    assert np.array_equal(d, expect), "tuple outside its partition range"
    a = np.stack([keys_in.astype(np.int64), np.asarray(pay_in).astype(np.int64)], 1)
    b = np.stack([keys_out.astype(np.int64), np.asarray(pay_out).astype(np.int64)], 1)
    a = a[np.lexsort((a[:, 1], a[:, 0]))]
    b = b[np.lexsort((b[:, 1], b[:, 0]))]
    assert np.array_equal(a, b), "output is not a permutation of the input tuples"


def table_size_for(base: int, occupancy: float = 0.5) -> int:
    """hash_table_size = ceil(base / HASH_TABLE_OCC) (src/partitioned_join.cuh:198)."""
    import math
    return int(math.ceil(base / occupancy))


def base_build_size(n_build: int, num_partitions: int, factor: float = 3.0) -> int:
    """base = ceil(f * |R| / P) (main.cu:867)."""
    import math
    return int(math.ceil(factor * n_build / num_partitions))
