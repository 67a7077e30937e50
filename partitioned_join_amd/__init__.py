"""MI355X-native radix-partitioned hash join.

Host-side mirror of the reference's partitioned-join interface
(src/partitioned_join.cuh, src/skewed_data_handler.cuh, main.cu:759-893) over
the C ABI of ``libpj.so`` (include/pj.h).  PyTorch supplies device memory and
streams only; all work runs in the hand-written gfx950 kernels.  There is no CPU
fallback: importing works without a GPU (for planning and introspection), every
compute call needs the HIP library and a device.

Tables are SoA like the reference: an int32 or int64 key tensor and an optional
int32 payload tensor (None = row id).  The join returns an (n, 2) int32 tensor of
(build payload, probe payload) pairs -- (build_idx, probe_idx) for row-id
payloads -- in unspecified order, like the reference's atomically claimed
output (src/hash_join_kernels.cuh:185-200).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Union

from . import _lib
from ._lib import PjError, PjPhaseTimes, PjPlan, PjTable

__all__ = [
    "PjError", "PjPlan", "plan_default", "workspace_size", "partitioned_join", "partitioned_join_async",
    "combined_histogram_pass", "histogram_pass", "exclusive_scan", "inclusive_scan",
    "partition_input_tables_pass_one", "partition_input_tables_pass_two", "handle_skewed_data",
    "shared_memory_hash_join", "project_reference", "hash_keys", "fastdiv", "tuple_keys", "tuple_payloads",
    "profile_enable", "profile_read", "profile_read_recent", "profile_select", "version", "lib",
]


def lib():
    return _lib.load()


def version() -> str:
    return lib().pj_version().decode()


def _torch():
    import torch
    return torch


def _stream(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _check_keys(keys, name: str):
    torch = _torch()
    if keys.dtype not in (torch.int32, torch.int64):
        raise TypeError(f"{name} keys must be int32 or int64, got {keys.dtype}")
    if not keys.is_contiguous():
        raise ValueError(f"{name} keys must be contiguous")


def _table(keys, payload=None, name="table") -> PjTable:
    torch = _torch()
    _check_keys(keys, name)
    if payload is not None:
        if payload.dtype != torch.int32 or not payload.is_contiguous() or payload.numel() != keys.numel():
            raise ValueError(f"{name} payload must be a contiguous int32 tensor of the key length")
    return PjTable(_ptr(keys), _ptr(payload), keys.numel(), keys.element_size(), 0)


def _size_table(n: int, key_bytes: int) -> PjTable:
    return PjTable(None, None, int(n), int(key_bytes), 0)


def _plan_struct(plan) -> Optional[PjPlan]:
    if plan is None or isinstance(plan, PjPlan):
        return plan
    p = PjPlan()
    p.passes = int(plan.get("passes", 2 if plan.get("bits2", 0) else 1))
    p.bits1 = int(plan["bits1"])
    p.bits2 = int(plan.get("bits2", 0))
    p.unique_build_keys = int(plan.get("unique_build_keys", 0))
    p.max_partition_size_factor = float(plan.get("max_partition_size_factor", 3.0))
    p.probe_slice = int(plan.get("probe_slice", 0))
    p.source = int(plan.get("source", 2))
    return p


def plan_default(n_build: int, n_probe: int, key_bytes: int = 4) -> PjPlan:
    """The plain histogram-driven plan (host only)."""
    plan = PjPlan()
    _lib.check(lib().pj_plan_default(ctypes.byref(_size_table(n_build, key_bytes)),
                                     ctypes.byref(_size_table(n_probe, key_bytes)), ctypes.byref(plan)))
    return plan


def workspace_size(n_build: int, n_probe: int, key_bytes: int = 4, plan=None) -> int:
    out = ctypes.c_size_t(0)
    p = _plan_struct(plan)
    _lib.check(lib().pj_workspace_size(ctypes.byref(_size_table(n_build, key_bytes)),
                                       ctypes.byref(_size_table(n_probe, key_bytes)),
                                       ctypes.byref(p) if p is not None else None, ctypes.byref(out)))
    return out.value


_WS = {}


def _workspace(nbytes: int, device):
    """Per-device cached workspace (256-byte aligned, grows on demand)."""
    torch = _torch()
    key = (device.type, device.index)
    ws = _WS.get(key)
    if ws is None or ws.numel() < nbytes:
        _WS.pop(key, None)
        ws = torch.empty(max(nbytes, 1 << 20), dtype=torch.uint8, device=device)
        _WS[key] = ws
    return ws


def partitioned_join(build_keys, probe_keys, build_payload=None, probe_payload=None, plan=None,
                     out_capacity: Optional[int] = None, stream=None, return_count: bool = False):
    """Radix-partitioned inner equi-join on the GPU (pj_partitioned_join).

    Returns an (n, 2) int32 device tensor of (build payload, probe payload)
    pairs.  ``out_capacity`` defaults to the probe size; if the join produces
    more pairs the call is repeated once with the exact count the library
    reports (PJ_ERR_CAPACITY)."""
    torch = _torch()
    bt = _table(build_keys, build_payload, "build")
    pt = _table(probe_keys, probe_payload, "probe")
    p = _plan_struct(plan)
    need = ctypes.c_size_t(0)
    _lib.check(lib().pj_workspace_size(ctypes.byref(bt), ctypes.byref(pt), ctypes.byref(p) if p is not None else None,
                                       ctypes.byref(need)))
    ws = _workspace(need.value, build_keys.device)
    cap = int(out_capacity) if out_capacity is not None else max(int(probe_keys.numel()), 1)
    count = ctypes.c_int64(0)
    for attempt in range(2):
        out = torch.empty((max(cap, 1), 2), dtype=torch.int32, device=build_keys.device)
        st = lib().pj_partitioned_join(ctypes.byref(bt), ctypes.byref(pt), ctypes.byref(p) if p is not None else None,
                                       _ptr(out), cap, ctypes.byref(count), _ptr(ws), ws.numel(), _stream(stream))
        if st == _lib.PJ_ERR_CAPACITY and out_capacity is None and attempt == 0:
            cap = int(count.value)
            continue
        _lib.check(st, count.value)
        break
    res = out[: count.value]
    return (res, count.value) if return_count else res


def partitioned_join_async(build_keys, probe_keys, out, build_payload=None, probe_payload=None, plan=None,
                           workspace=None, stream=None):
    """Stream-ordered join into ``out`` ((cap, 2) int32).  Returns a device int64
    tensor holding the pair count (no host synchronisation)."""
    torch = _torch()
    bt = _table(build_keys, build_payload, "build")
    pt = _table(probe_keys, probe_payload, "probe")
    p = _plan_struct(plan)
    need = ctypes.c_size_t(0)
    _lib.check(lib().pj_workspace_size(ctypes.byref(bt), ctypes.byref(pt), ctypes.byref(p) if p is not None else None,
                                       ctypes.byref(need)))
    ws = workspace if workspace is not None else _workspace(need.value, build_keys.device)
    d_count = torch.zeros(1, dtype=torch.int64, device=build_keys.device)
    _lib.check(lib().pj_partitioned_join_async(ctypes.byref(bt), ctypes.byref(pt),
                                               ctypes.byref(p) if p is not None else None, _ptr(out),
                                               out.shape[0], _ptr(d_count), _ptr(ws), ws.numel(), _stream(stream)))
    return d_count


def partitioned_join_noalloc(build_keys, probe_keys, out, d_count, build_payload=None, probe_payload=None, plan=None,
                             stream=None) -> None:
    """pj_partitioned_join_noalloc: the join into ``out`` ((cap, 2) int32) with
    the pair count written to ``d_count`` (device int64, one element), using the
    library's static device scratch instead of a workspace; allocates nothing.
    Calls that overlap on one GPU race (they share the scratch)."""
    bt = _table(build_keys, build_payload, "build")
    pt = _table(probe_keys, probe_payload, "probe")
    p = _plan_struct(plan)
    _lib.check(lib().pj_partitioned_join_noalloc(ctypes.byref(bt), ctypes.byref(pt),
                                                 ctypes.byref(p) if p is not None else None, _ptr(out), out.shape[0],
                                                 _ptr(d_count), _stream(stream)))


def project_reference(build_keys, pairs, stream=None):
    """(build key, probe payload) pairs: the reference's output form."""
    torch = _torch()
    out = torch.empty((pairs.shape[0], 2), dtype=build_keys.dtype, device=build_keys.device)
    _lib.check(lib().pj_project_reference(ctypes.byref(_table(build_keys, None, "build")), _ptr(pairs),
                                          pairs.shape[0], _ptr(out), _stream(stream)))
    return out


def gather(pairs, build_col=None, probe_col=None, stream=None):
    """Payload columns of the joined rows (cuDF gather map): returns
    (build_col[pairs[:, 0]], probe_col[pairs[:, 1]]), None for a column not
    given.  Columns are int32/int64 (or float32/float64) device tensors."""
    torch = _torch()
    n = pairs.shape[0]

    def prep(col):
        if col is None:
            return None, 4, None
        if col.element_size() not in (4, 8) or not col.is_contiguous():
            raise ValueError("gathered columns must be contiguous with 4- or 8-byte elements")
        return col, col.element_size(), torch.empty(n, dtype=col.dtype, device=pairs.device)

    b, bb, ob = prep(build_col)
    p, pb, op = prep(probe_col)
    _lib.check(lib().pj_gather(_ptr(pairs), n, _ptr(b) if b is not None else None, bb,
                               _ptr(p) if p is not None else None, pb,
                               _ptr(ob) if ob is not None else None, _ptr(op) if op is not None else None,
                               _stream(stream)))
    return ob, op


# --------------------------------------------------------------------------- phases
def combined_histogram_pass(keys, total_bits: int, bits2: int, stream=None):
    """(fine histogram of key & (2^total_bits-1), pass-1 histogram) as int32
    tensors holding uint32 counts (combined_histogram_pass,
    src/partitioned_join.cuh:22-48)."""
    torch = _torch()
    fine = torch.empty(1 << total_bits, dtype=torch.int32, device=keys.device)
    p1 = torch.empty(1 << (total_bits - bits2), dtype=torch.int32, device=keys.device)
    nb = lib().pj_histogram_workspace_size(total_bits)
    ws = _workspace(nb, keys.device)
    _lib.check(lib().pj_combined_histogram(ctypes.byref(_table(keys)), total_bits, bits2, _ptr(fine), _ptr(p1),
                                           _ptr(ws), ws.numel(), _stream(stream)))
    return fine, p1


def histogram_pass(keys, shift: int, bits: int, stream=None):
    """Histogram of (key >> shift) & (2^bits - 1) (histogram_pass_one,
    src/partition_kernels.cuh:140-194)."""
    torch = _torch()
    h = torch.empty(1 << bits, dtype=torch.int32, device=keys.device)
    _lib.check(lib().pj_histogram(ctypes.byref(_table(keys)), shift, bits, _ptr(h), None, 0, _stream(stream)))
    return h


def _scan(x, inclusive: bool, stream=None):
    torch = _torch()
    out = torch.empty_like(x)
    ws = _workspace(lib().pj_scan_workspace_size(x.numel()), x.device)
    fn = lib().pj_inclusive_scan if inclusive else lib().pj_exclusive_scan
    _lib.check(fn(_ptr(x), _ptr(out), x.numel(), _ptr(ws), ws.numel(), _stream(stream)))
    return out


def exclusive_scan(x, stream=None):
    """cub::DeviceScan::ExclusiveSum over uint32 (held in an int32 tensor)."""
    return _scan(x, False, stream)


def inclusive_scan(x, stream=None):
    """cub::DeviceScan::InclusiveSum over uint32 (held in an int32 tensor)."""
    return _scan(x, True, stream)


def _tuple_cols(key_bytes: int) -> int:
    return 2 if key_bytes == 4 else 4


def tuple_keys(tuples, key_bytes: int):
    """Keys of pj tuples ((n, 2) int32 for 4-byte keys, (n, 4) int32 for 8-byte keys)."""
    if key_bytes == 4:
        return tuples[:, 0].contiguous()
    return tuples.contiguous().view(-1, 2, 2)[:, 0, :].contiguous().view(-1).view(_torch().int64)


def tuple_payloads(tuples, key_bytes: int):
    return (tuples[:, 1] if key_bytes == 4 else tuples[:, 2]).contiguous()


def partition_input_tables_pass_one(keys, payload=None, bits: int = 8, shift: int = 0, stream=None):
    """One radix pass over a table: returns (pj tuples, offsets[2^bits + 1])
    (partition_input_tables_pass_one, src/partitioned_join.cuh:50-107)."""
    torch = _torch()
    t = _table(keys, payload)
    kb = keys.element_size()
    out = torch.empty((max(keys.numel(), 1), _tuple_cols(kb)), dtype=torch.int32, device=keys.device)
    off = torch.empty((1 << bits) + 1, dtype=torch.int32, device=keys.device)
    ws = _workspace(lib().pj_partition_workspace_size(ctypes.byref(t), bits), keys.device)
    _lib.check(lib().pj_partition_pass(ctypes.byref(t), shift, bits, _ptr(out), _ptr(off), _ptr(ws), ws.numel(),
                                       _stream(stream)))
    return out[: keys.numel()], off


def partition_input_tables_pass_two(tuples, key_bytes: int, bits1: int, bits2: int, offsets1, stream=None):
    """Second pass: split every pass-1 partition by key & (2^bits2 - 1)
    (partition_input_tables_pass_two, src/partitioned_join.cuh:109-175)."""
    torch = _torch()
    n = tuples.shape[0]
    out = torch.empty((max(n, 1), _tuple_cols(key_bytes)), dtype=torch.int32, device=tuples.device)
    off2 = torch.empty((1 << (bits1 + bits2)) + 1, dtype=torch.int32, device=tuples.device)
    ws = _workspace(lib().pj_partition2_workspace_size(n, bits1, bits2), tuples.device)
    _lib.check(lib().pj_partition_pass2(_ptr(tuples), n, key_bytes, bits1, bits2, _ptr(offsets1), _ptr(out),
                                        _ptr(off2), _ptr(ws), ws.numel(), _stream(stream)))
    return out[:n], off2


def handle_skewed_data(offsets, total: int, num_partitions: int, base: int, max_blocks: Optional[int] = None,
                       stream=None):
    """Block assignment of handle_skewed_data (src/skewed_data_handler.cuh:58-91):
    (blocks_scanned inclusive, partition_index[block], total_blocks tensor)."""
    torch = _torch()
    if max_blocks is None:
        max_blocks = num_partitions + (int(total) + base - 1) // base
    scanned = torch.empty(num_partitions, dtype=torch.int32, device=offsets.device)
    pidx = torch.empty(max(max_blocks, 1), dtype=torch.int32, device=offsets.device)
    tot = torch.empty(1, dtype=torch.int32, device=offsets.device)
    ws = _workspace(lib().pj_assign_blocks_workspace_size(num_partitions), offsets.device)
    _lib.check(lib().pj_assign_blocks(_ptr(offsets), int(total), num_partitions, base, _ptr(scanned), _ptr(pidx),
                                      max_blocks, _ptr(tot), _ptr(ws), ws.numel(), _stream(stream)))
    return scanned, pidx, tot


def shared_memory_hash_join(build_tuples, build_offsets, probe_tuples, probe_offsets, key_bytes: int,
                            total_bits: int, base: int, probe_slice: int = 0, unique_build_keys: bool = False,
                            out_capacity: Optional[int] = None, stream=None):
    """Join of partitioned pj tuples (shared_memory_hash_join + hash_join,
    src/partitioned_join.cuh:177-251).  Returns (pairs tensor, count)."""
    torch = _torch()
    cap = int(out_capacity) if out_capacity is not None else max(int(probe_tuples.shape[0]), 1)
    out = torch.empty((max(cap, 1), 2), dtype=torch.int32, device=build_tuples.device)
    d_count = torch.zeros(1, dtype=torch.int64, device=build_tuples.device)
    ws = _workspace(lib().pj_join_workspace_size(total_bits), build_tuples.device)
    _lib.check(lib().pj_join_partitions(_ptr(build_tuples), _ptr(build_offsets), _ptr(probe_tuples),
                                        _ptr(probe_offsets), key_bytes, total_bits, base, probe_slice,
                                        int(bool(unique_build_keys)), _ptr(out), cap, _ptr(d_count), _ptr(ws),
                                        ws.numel(), _stream(stream)))
    n = int(d_count.item())
    return out[: min(n, cap)], n


# --------------------------------------------------------------------------- primitives
def hash_keys(keys, stream=None):
    """MurmurHash3_x86_32 (seed 0) of every key, as int32 holding uint32."""
    torch = _torch()
    _check_keys(keys, "hash")
    out = torch.empty(keys.numel(), dtype=torch.int32, device=keys.device)
    _lib.check(lib().pj_hash_keys(_ptr(keys), keys.numel(), keys.element_size(), _ptr(out), _stream(stream)))
    return out


def fastdiv(n, divisor: int, stream=None):
    """(n / d, n % d) with int_fastdiv on the device."""
    torch = _torch()
    q = torch.empty_like(n)
    r = torch.empty_like(n)
    _lib.check(lib().pj_fastdiv(_ptr(n), n.numel(), divisor, _ptr(q), _ptr(r), _stream(stream)))
    return q, r


# --------------------------------------------------------------------------- profiling
def profile_enable(on: bool = True) -> None:
    _lib.check(lib().pj_profile_enable(int(bool(on))))


def profile_read() -> dict:
    t = PjPhaseTimes()
    _lib.check(lib().pj_profile_read(ctypes.byref(t)))
    return t.as_dict()


def profile_select(marks: int = 0xFF) -> None:
    """The phase boundaries later profiled joins record (bit i: the event
    before phase i, bit 7 the last); phases without both events read 0."""
    _lib.check(lib().pj_profile_select(int(marks)))


def profile_read_recent(n: int) -> list:
    """Phase times of the last n profiled joins on this thread (at most 128),
    oldest first, after one wait for the newest."""
    times = (PjPhaseTimes * max(1, n))()
    got = ctypes.c_int32(0)
    _lib.check(lib().pj_profile_read_recent(times, int(n), ctypes.byref(got)))
    return [times[i].as_dict() for i in range(got.value)]


from . import generators  # noqa: E402,F401
