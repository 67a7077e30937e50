"""ctypes binding of libpj.so (the C ABI declared in include/pj.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpj.so")

c_i32, c_i64, c_u32, c_u64, c_dbl, c_vp, c_size = (
    ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_void_p,
    ctypes.c_size_t)

PJ_OK, PJ_ERR_INVALID_ARG, PJ_ERR_CAPACITY, PJ_ERR_WORKSPACE, PJ_ERR_HIP, PJ_ERR_UNSUPPORTED = range(6)
STATUS_NAMES = {0: "PJ_OK", 1: "PJ_ERR_INVALID_ARG", 2: "PJ_ERR_CAPACITY", 3: "PJ_ERR_WORKSPACE",
                4: "PJ_ERR_HIP", 5: "PJ_ERR_UNSUPPORTED"}


class PjTable(ctypes.Structure):
    _fields_ = [("keys", c_vp), ("payload", c_vp), ("n", c_i64), ("key_bytes", c_i32), ("reserved", c_i32)]


class PjPlan(ctypes.Structure):
    _fields_ = [("passes", c_i32), ("bits1", c_i32), ("bits2", c_i32), ("unique_build_keys", c_i32),
                ("max_partition_size_factor", c_dbl), ("probe_slice", c_i64), ("source", c_i32),
                ("reserved", c_i32 * 7)]

    def as_dict(self):
        return {"passes": self.passes, "bits1": self.bits1, "bits2": self.bits2,
                "unique_build_keys": self.unique_build_keys,
                "max_partition_size_factor": self.max_partition_size_factor,
                "probe_slice": self.probe_slice, "source": self.source}


class PjPhaseTimes(ctypes.Structure):
    _fields_ = [("histogram_ms", ctypes.c_float), ("offsets_ms", ctypes.c_float), ("pass1_ms", ctypes.c_float),
                ("pass2_ms", ctypes.c_float), ("skew_ms", ctypes.c_float), ("join_ms", ctypes.c_float),
                ("compact_ms", ctypes.c_float), ("total_ms", ctypes.c_float), ("valid", c_i32),
                ("reserved", c_i32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


P_TABLE = ctypes.POINTER(PjTable)
P_PLAN = ctypes.POINTER(PjPlan)

# name -> (restype, argtypes).  Every symbol declared in include/pj.h.
SIGNATURES = {
    "pj_version": (ctypes.c_char_p, []),
    "pj_last_error": (ctypes.c_char_p, []),
    "pj_plan_default": (c_i32, [P_TABLE, P_TABLE, P_PLAN]),
    "pj_workspace_size": (c_i32, [P_TABLE, P_TABLE, P_PLAN, ctypes.POINTER(c_size)]),
    "pj_partitioned_join": (c_i32, [P_TABLE, P_TABLE, P_PLAN, c_vp, c_i64, ctypes.POINTER(c_i64), c_vp, c_size,
                                    c_vp]),
    "pj_partitioned_join_async": (c_i32, [P_TABLE, P_TABLE, P_PLAN, c_vp, c_i64, c_vp, c_vp, c_size, c_vp]),
    "pj_partitioned_join_noalloc": (c_i32, [P_TABLE, P_TABLE, P_PLAN, c_vp, c_i64, c_vp, c_vp]),
    "pj_project_reference": (c_i32, [P_TABLE, c_vp, c_i64, c_vp, c_vp]),
    "pj_gather": (c_i32, [c_vp, c_i64, c_vp, c_i32, c_vp, c_i32, c_vp, c_vp, c_vp]),
    "pj_combined_histogram": (c_i32, [P_TABLE, c_i32, c_i32, c_vp, c_vp, c_vp, c_size, c_vp]),
    "pj_histogram_workspace_size": (c_size, [c_i32]),
    "pj_histogram": (c_i32, [P_TABLE, c_i32, c_i32, c_vp, c_vp, c_size, c_vp]),
    "pj_exclusive_scan": (c_i32, [c_vp, c_vp, c_i64, c_vp, c_size, c_vp]),
    "pj_inclusive_scan": (c_i32, [c_vp, c_vp, c_i64, c_vp, c_size, c_vp]),
    "pj_scan_workspace_size": (c_size, [c_i64]),
    "pj_partition_pass": (c_i32, [P_TABLE, c_i32, c_i32, c_vp, c_vp, c_vp, c_size, c_vp]),
    "pj_partition_workspace_size": (c_size, [P_TABLE, c_i32]),
    "pj_partition_pass2": (c_i32, [c_vp, c_i64, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "pj_partition2_workspace_size": (c_size, [c_i64, c_i32, c_i32]),
    "pj_assign_blocks": (c_i32, [c_vp, c_i64, c_i32, c_i32, c_vp, c_vp, c_i64, c_vp, c_vp, c_size, c_vp]),
    "pj_assign_blocks_workspace_size": (c_size, [c_i32]),
    "pj_join_partitions": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_i32, c_i32, c_i32, c_i64, c_i32, c_vp, c_i64, c_vp,
                                   c_vp, c_size, c_vp]),
    "pj_join_workspace_size": (c_size, [c_i32]),
    "pj_hash_keys": (c_i32, [c_vp, c_i64, c_i32, c_vp, c_vp]),
    "pj_fastdiv": (c_i32, [c_vp, c_i64, c_i32, c_vp, c_vp, c_vp]),
    "pj_gen_build_unique_i32": (c_i32, [c_vp, c_i64, c_i64, c_u64, c_vp]),
    "pj_gen_build_dup_i32": (c_i32, [c_vp, c_i64, c_i64, c_i64, c_u64, c_u64, c_vp]),
    "pj_gen_probe_i32": (c_i32, [c_vp, c_i64, c_i64, c_i64, c_u64, c_u64, c_dbl, c_i64, c_i64, c_vp]),
    "pj_zipf_lut": (c_i32, [c_vp, c_i64, c_dbl]),
    "pj_gen_probe_zipf_i32": (c_i32, [c_vp, c_i64, c_vp, c_i64, c_u64, c_u64, c_vp]),
    "pj_gen_build_i64": (c_i32, [c_vp, c_i64, c_u64, c_vp]),
    "pj_gen_probe_i64": (c_i32, [c_vp, c_i64, c_i64, c_u64, c_u64, c_dbl, c_vp]),
    "pj_gen_probe_zipf_i64": (c_i32, [c_vp, c_i64, c_vp, c_i64, c_u64, c_u64, c_u64, c_vp]),
    "pj_gen_row_ids": (c_i32, [c_vp, c_i64, c_vp]),
    "pj_profile_enable": (c_i32, [c_i32]),
    "pj_profile_read": (c_i32, [ctypes.POINTER(PjPhaseTimes)]),
    "pj_profile_read_recent": (c_i32, [ctypes.POINTER(PjPhaseTimes), c_i32, ctypes.POINTER(c_i32)]),
    "pj_profile_select": (c_i32, [c_u32]),
    "pj_planner_param_count": (c_size, []),
    "pj_join_chunk_capacity": (c_i32, [c_i32]),
    "pj_planner_scratch_bytes": (c_size, [c_i32]),
    "pj_planner_sketch": (c_i32, [P_TABLE, P_TABLE, c_vp, c_vp, c_vp, c_vp]),
    "pj_planner_candidate_counts": (c_i32, [P_TABLE, c_vp, c_vp, c_vp, c_vp]),
    "pj_planner_forward": (c_i32, [c_vp, c_vp, c_i32, c_vp, c_vp, c_size, c_vp]),
    "pj_planner_train_step": (c_i32, [c_vp, c_vp, c_vp, c_i32, ctypes.c_float, c_vp, c_vp, c_vp, c_i32, c_vp, c_vp,
                                      c_size, c_vp]),
    "pj_plan_decode": (c_i32, [c_vp, c_vp, c_vp, P_TABLE, P_TABLE, P_PLAN, ctypes.POINTER(c_i32)]),
    "pj_plan_learned_workspace_size": (c_size, []),
    "pj_plan_learned": (c_i32, [P_TABLE, P_TABLE, c_vp, P_PLAN, ctypes.POINTER(c_i32), c_vp, c_size, c_vp]),
    "pj_batch_workspace_size": (c_i32, [P_TABLE, P_TABLE, P_PLAN, c_i32, c_i32, ctypes.POINTER(c_size)]),
    "pj_batch_create": (c_i32, [P_TABLE, P_TABLE, P_PLAN, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_size,
                                ctypes.POINTER(c_vp)]),
    "pj_batch_launch": (c_i32, [c_vp, c_vp]),
    "pj_batch_destroy": (c_i32, [c_vp]),
}

_LIB = None


def hip_runtime_path() -> str:
    """The HIP runtime libpj.so binds to: PyTorch's bundled libamdhip64 when
    torch is a ROCm build (so the process has exactly one runtime), else the
    system ROCm one."""
    try:
        import torch
        p = os.path.join(os.path.dirname(torch.__file__), "lib", "libamdhip64.so")
        if os.path.exists(p):
            return p
    except ImportError:
        pass
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "lib", "libamdhip64.so")


def load() -> ctypes.CDLL:
    """Load libpj.so; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libpj.so not found at {LIB_PATH}: build it with `make -C partitioned_join_amd` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        # libpj.so leaves the HIP runtime symbols undefined; make the runtime the
        # process uses globally visible before loading it.
        ctypes.CDLL(hip_runtime_path(), mode=ctypes.RTLD_GLOBAL)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


class PjError(RuntimeError):
    def __init__(self, status: int, message: str, count: int = -1):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.count = count


def check(status: int, count: int = -1) -> None:
    if status != PJ_OK:
        msg = load().pj_last_error().decode(errors="replace")
        raise PjError(status, msg, count)
