"""Seeded synthetic inputs generated on the device (pj_gen_* in libpj.so).

Definitions (shared bit-for-bit with the CPU restatement used by the tests):
SplitMix64 counters rand64(seed, i) = mix64(seed + (i + 1) * 0x9e3779b97f4a7c15),
bounded(r, n) = (hi32(r) * n) >> 32, and perm_n(i), a keyed bijection of
[0, 2^k) (four rounds of odd multiply-add and xorshift) cycle-walked onto
[0, n).  The workloads follow the reference generator's semantics
(src/generate_input_tables.cuh:116-141: unique lottery, selectivity) and Zipf
sampler (src/generate_zipf_input.cuh:50-127).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _t():
    import torch
    return torch


def _s(stream=None):
    torch = _t()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def build_unique_i32(n: int, domain: int, perm_seed: int, device="cuda", stream=None):
    torch = _t()
    out = torch.empty(n, dtype=torch.int32, device=device)
    _lib.check(_lib.load().pj_gen_build_unique_i32(out.data_ptr(), n, domain, perm_seed, _s(stream)))
    return out


def build_dup_i32(n: int, m: int, domain: int, perm_seed: int, seed: int, device="cuda", stream=None):
    torch = _t()
    out = torch.empty(n, dtype=torch.int32, device=device)
    _lib.check(_lib.load().pj_gen_build_dup_i32(out.data_ptr(), n, m, domain, perm_seed, seed, _s(stream)))
    return out


def probe_i32(n: int, domain: int, n_build: int, perm_seed: int, seed: int, sel: float, nonmatch_base: int = 0,
              nonmatch_range: int = 0, device="cuda", stream=None):
    torch = _t()
    out = torch.empty(n, dtype=torch.int32, device=device)
    _lib.check(_lib.load().pj_gen_probe_i32(out.data_ptr(), n, domain, n_build, perm_seed, seed, sel, nonmatch_base,
                                            nonmatch_range, _s(stream)))
    return out


def zipf_lut(n: int, theta: float) -> np.ndarray:
    lut = np.empty(n, dtype=np.float64)
    _lib.check(_lib.load().pj_zipf_lut(lut.ctypes.data_as(ctypes.c_void_p), n, theta))
    return lut


def probe_zipf_i32(count: int, n: int, theta: float, rank_perm_seed: int, seed: int, device="cuda", stream=None,
                   lut=None):
    torch = _t()
    if lut is None:
        lut = zipf_lut(n, theta)
    d_lut = torch.from_numpy(lut).to(device)
    out = torch.empty(count, dtype=torch.int32, device=device)
    _lib.check(_lib.load().pj_gen_probe_zipf_i32(out.data_ptr(), count, d_lut.data_ptr(), n, rank_perm_seed, seed,
                                                 _s(stream)))
    return out


def build_i64(n: int, key_seed: int, device="cuda", stream=None):
    torch = _t()
    out = torch.empty(n, dtype=torch.int64, device=device)
    _lib.check(_lib.load().pj_gen_build_i64(out.data_ptr(), n, key_seed, _s(stream)))
    return out


def probe_i64(n: int, n_build: int, key_seed: int, seed: int, sel: float, device="cuda", stream=None):
    torch = _t()
    out = torch.empty(n, dtype=torch.int64, device=device)
    _lib.check(_lib.load().pj_gen_probe_i64(out.data_ptr(), n, n_build, key_seed, seed, sel, _s(stream)))
    return out


def probe_zipf_i64(count: int, n_build: int, theta: float, key_seed: int, rank_perm_seed: int, seed: int,
                   device="cuda", stream=None, lut=None):
    torch = _t()
    if lut is None:
        lut = zipf_lut(n_build, theta)
    d_lut = torch.from_numpy(lut).to(device)
    out = torch.empty(count, dtype=torch.int64, device=device)
    _lib.check(_lib.load().pj_gen_probe_zipf_i64(out.data_ptr(), count, d_lut.data_ptr(), n_build, key_seed,
                                                 rank_perm_seed, seed, _s(stream)))
    return out


def row_ids(n: int, device="cuda", stream=None):
    torch = _t()
    out = torch.empty(n, dtype=torch.int32, device=device)
    _lib.check(_lib.load().pj_gen_row_ids(out.data_ptr(), n, _s(stream)))
    return out


# Named workloads (sizes of the benchmark configurations).
SEEDS = {"c1": (0xC1B, 0xC1F), "c2": (0xC2B, 0xC2F), "c3": (0xC3B, 0xC3F), "c3rank": 0xC3A}


def workload_uniform(n_build: int, n_probe: int, sel: float, nonmatch: bool, seeds, device="cuda"):
    """Build = perm of [0, n_build); probe = a build key with probability sel,
    otherwise (nonmatch) n_build + U(0, n_build)."""
    bseed, pseed = seeds
    bk = build_unique_i32(n_build, n_build, bseed, device)
    if sel >= 1.0 or not nonmatch:
        pk = probe_i32(n_probe, n_build, n_build, bseed, pseed, 1.0, device=device)
    else:
        pk = probe_i32(n_probe, n_build, n_build, bseed, pseed, sel, n_build, n_build, device=device)
    return bk, pk


def workload_config1(device="cuda"):
    """1M unique keys (perm of [0, 2^20)) joined with 4M probes drawn from them."""
    return workload_uniform(1 << 20, 1 << 22, 1.0, False, SEEDS["c1"], device)


def workload_config2(device="cuda", scale_shift: int = 0):
    """64M unique keys (perm of [0, 2^26)) joined with 256M probes, half of them
    build keys and half drawn from [2^26, 2^27)."""
    nb, npb = (1 << 26) >> scale_shift, (1 << 28) >> scale_shift
    return workload_uniform(nb, npb, 0.5, True, SEEDS["c2"], device)


def workload_config3(theta: float, device="cuda", scale_shift: int = 0, lut=None):
    """16M unique keys joined with 256M Zipf(theta) probes over them."""
    nb, npb = (1 << 24) >> scale_shift, (1 << 28) >> scale_shift
    bseed, pseed = SEEDS["c3"]
    bk = build_unique_i32(nb, nb, bseed, device)
    pk = probe_zipf_i32(npb, nb, theta, SEEDS["c3rank"], pseed, device, lut=lut)
    return bk, pk
