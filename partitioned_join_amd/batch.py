"""Batched joins: many independent joins of uneven sizes replayed as one HIP
graph (pj_batch_* in libpj.so), with an autotuner for the number of parallel
chains, the radix fan-out and the join's work-item tile.

The joins are spread longest-first over parallel chains that are captured
once into one graph (pj_batch_create). Each chain reuses one workspace region
for its joins. One pj_batch_launch then replays everything. The autotuner
times whole-batch replays over:
- chain counts;
- a fan-out offset (-1, 0 or +1 total radix bits) applied to every join's
  plan;
- the join's work-item tile: the build chunk factor and the probe slice.

It keeps the fastest combination. Plans come from the plain planner, or from
the learned planner per join, with its plain-plan fallback.
"""
from __future__ import annotations

import ctypes
import math
import time
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import PjPlan, PjTable


def _t():
    import torch
    return torch


def _s(stream=None):
    torch = _t()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _L():
    return _lib.load()


def _table(keys, payload=None) -> PjTable:
    return PjTable(keys.data_ptr(), None if payload is None else payload.data_ptr(), keys.numel(),
                   keys.element_size(), 0)


def shift_plan(plan: PjPlan, delta: int, n_build: int, key_bytes: int) -> PjPlan:
    """The plan with `delta` more (or fewer) total radix bits, taken from the
    second pass (or the only one); kept within 1..12 bits per pass, 16 in
    total, and above the fan-out whose build chunks would not fit the LDS
    table."""
    from .planner import chunk_capacity
    p = PjPlan()
    ctypes.memmove(ctypes.byref(p), ctypes.byref(plan), ctypes.sizeof(PjPlan))
    if delta == 0:
        return p
    cap = chunk_capacity(key_bytes)
    if p.passes == 2:
        b2 = min(12, max(1, p.bits2 + delta))
        if p.bits1 + b2 > 16:
            b2 = 16 - p.bits1
        p.bits2 = b2
    else:
        p.bits1 = min(12, max(1, p.bits1 + delta))
    if p.max_partition_size_factor * n_build / float(1 << (p.bits1 + p.bits2)) > cap:
        ctypes.memmove(ctypes.byref(p), ctypes.byref(plan), ctypes.sizeof(PjPlan))
    p.source = 2
    return p


class BatchJoin:
    """One captured graph over a list of joins (device key tensors)."""

    def __init__(self, builds: Sequence, probes: Sequence, plans: Optional[Sequence[PjPlan]] = None,
                 chains: int = 8, out_caps: Optional[Sequence[int]] = None, build_payloads=None,
                 probe_payloads=None):
        torch = _t()
        n = len(builds)
        assert n == len(probes) and n > 0
        dev = builds[0].device
        self.builds, self.probes = list(builds), list(probes)
        self.keep = [build_payloads, probe_payloads]
        tb = (PjTable * n)(*[_table(builds[j], None if build_payloads is None else build_payloads[j])
                             for j in range(n)])
        tp = (PjTable * n)(*[_table(probes[j], None if probe_payloads is None else probe_payloads[j])
                             for j in range(n)])
        self.tb, self.tp = tb, tp
        self.plans = None
        if plans is not None:
            self.plans = (PjPlan * n)(*plans)
        caps = [int(c) for c in (out_caps if out_caps is not None else [max(1, p.numel()) for p in probes])]
        offs = np.zeros(n, dtype=np.int64)
        offs[1:] = np.cumsum(caps)[:-1]
        self.offsets = offs
        self.caps = np.asarray(caps, dtype=np.int64)
        self.out = torch.empty((int(sum(caps)), 2), dtype=torch.int32, device=dev)
        self.d_counts = torch.zeros(n, dtype=torch.int64, device=dev)
        need = ctypes.c_size_t(0)
        _lib.check(_L().pj_batch_workspace_size(tb, tp, self.plans, n, chains, ctypes.byref(need)))
        self.ws = torch.empty(max(int(need.value), 256), dtype=torch.uint8, device=dev)
        self.handle = ctypes.c_void_p(0)
        _lib.check(_L().pj_batch_create(tb, tp, self.plans, n, chains, self.out.data_ptr(),
                                        self.offsets.ctypes.data, self.caps.ctypes.data, self.d_counts.data_ptr(),
                                        self.ws.data_ptr(), self.ws.numel(), ctypes.byref(self.handle)))
        self.count = n
        self.chains = chains
        self.tuples = sum(int(b.numel()) + int(p.numel()) for b, p in zip(builds, probes))

    def launch(self, stream=None) -> None:
        _lib.check(_L().pj_batch_launch(self.handle, _s(stream)))

    def counts(self) -> np.ndarray:
        return self.d_counts.cpu().numpy()

    def pairs(self, j: int, counts: Optional[np.ndarray] = None):
        c = int((counts if counts is not None else self.counts())[j])
        if c > self.caps[j]:
            raise _lib.PjError(_lib.PJ_ERR_CAPACITY, f"join {j}: {c} pairs exceed its capacity {self.caps[j]}", c)
        o = int(self.offsets[j])
        return self.out[o:o + c]

    def close(self) -> None:
        if self.handle:
            _L().pj_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def time(self, reps: int = 3) -> float:
        """Milliseconds per replay (warm-up replay excluded)."""
        torch = _t()
        self.launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            self.launch()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1000.0 / reps


def batch_plans(builds, probes, planner: str = "plain") -> List[PjPlan]:
    """Per-join plans: the plain plan, or the learned plan (with its fallback)."""
    from . import plan_default
    plans = []
    w = None
    fallbacks = 0
    if planner == "learned":
        from . import planner as P
        w = P.load_weights(device=builds[0].device)
    for b, p in zip(builds, probes):
        if w is not None:
            from . import planner as P
            plan, fb = P.plan_learned(b, p, w)
            fallbacks += int(fb)
        else:
            plan = plan_default(b.numel(), p.numel(), b.element_size())
        plans.append(plan)
    return plans


def tile_plan(plan: PjPlan, factor: Optional[float], probe_slice: Optional[int]) -> PjPlan:
    """The plan with the join's work-item tile changed: the build chunk factor
    f (chunk = ceil(f |R| / P), within the LDS table) and the probe slice
    (0 = automatic); None keeps the plan's own."""
    p = PjPlan()
    ctypes.memmove(ctypes.byref(p), ctypes.byref(plan), ctypes.sizeof(PjPlan))
    if factor is not None:
        p.max_partition_size_factor = float(factor)
    if probe_slice is not None:
        p.probe_slice = int(probe_slice)
    return p


def autotune(builds, probes, base_plans: Sequence[PjPlan], chain_options=(4, 8, 16, 32), bit_offsets=(-1, 0, 1),
             tiles=((None, None),), reps: int = 3, log=None):
    """Times the batch under every (chains, fan-out offset, join tile) and
    returns (best BatchJoin, best config, all timings in ms).  A tile is
    (build chunk factor, probe slice), None keeping each plan's own."""
    timings = {}
    best = None
    for delta in bit_offsets:
        shifted = [shift_plan(p, delta, b.numel(), b.element_size()) for p, b in zip(base_plans, builds)]
        for tile in tiles:
            plans = [tile_plan(p, *tile) for p in shifted]
            for chains in chain_options:
                bj = BatchJoin(builds, probes, plans, chains=chains)
                ms = bj.time(reps)
                bj.close()
                key = (chains, delta) if tile == (None, None) else (chains, delta, tile)
                timings[key] = ms
                if log:
                    log(f"chains={chains} fanout{delta:+d} tile={tile}: {ms:.2f} ms")
                if best is None or ms < best[0]:
                    best = (ms, plans, chains, key)
    # Every timed batch is destroyed before the chosen configuration is built
    # again: no returned graph lives through the destruction of another.
    return BatchJoin(builds, probes, best[1], chains=best[2]), best[3], timings


def make_batch_workload(count: int = 512, seed: int = 0xBA7C, device="cuda"):
    """Batched uneven joins: |R| log-uniform in [10K, 4M], |S| 1-8x |R|,
    uniform 64-bit random build keys (a collision among 4M draws has
    probability below 1e-6), probe keys Zipf(theta) over the build keys with
    theta uniform in [0, 1.2], int32 row-id payloads."""
    torch = _t()
    L = _L()
    s = _s()
    builds, probes, specs = [], [], []
    rng = np.random.default_rng(seed)
    for j in range(count):
        nb = int(round(math.exp(rng.uniform(math.log(1e4), math.log(4e6)))))
        ns = int(round(nb * rng.uniform(1.0, 8.0)))
        theta = float(rng.uniform(0.0, 1.2))
        k0, k1, k2 = (int(x) for x in rng.integers(1, 2**62, 3))
        bk = torch.empty(nb, dtype=torch.int64, device=device)
        _lib.check(L.pj_gen_build_i64(bk.data_ptr(), nb, k0, s))
        w = torch.arange(1, nb + 1, dtype=torch.float64, device=device).pow_(-theta)
        lut = torch.cumsum(w, 0)
        lut /= lut[-1].clone()
        pk = torch.empty(ns, dtype=torch.int64, device=device)
        _lib.check(L.pj_gen_probe_zipf_i64(pk.data_ptr(), ns, lut.data_ptr(), nb, k0, k1, k2, s))
        del w, lut
        builds.append(bk)
        probes.append(pk)
        specs.append({"n_build": nb, "n_probe": ns, "theta": theta})
    torch.cuda.synchronize()
    return builds, probes, specs
