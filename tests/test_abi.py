"""The C ABI library loads and exports every symbol include/pj.h declares, and
its host-only logic (plans, workspace sizes, argument checks, the Zipf table)
behaves as documented.  No device work is issued here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "pj.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pj_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(pj):
    from partitioned_join_amd import _lib
    names = declared_functions()
    assert len(names) >= 30
    L = pj.lib()
    for n in names:
        assert hasattr(L, n), f"{n} not exported by libpj.so"
        assert n in _lib.SIGNATURES, f"{n} missing from the ctypes binding"
    assert set(_lib.SIGNATURES) == set(names)


def test_version(pj):
    assert "gfx950" in pj.version()


def test_plain_plans(pj):
    # partitions of at most 2048 build rows on average for 4-byte keys, 1024 for 8-byte keys
    p1 = pj.plan_default(1 << 20, 1 << 22)
    assert (p1.passes, p1.bits1, p1.bits2) == (1, 9, 0)
    p2 = pj.plan_default(1 << 26, 1 << 28)
    assert (p2.passes, p2.bits1, p2.bits2) == (2, 8, 7)
    p3 = pj.plan_default(1 << 24, 1 << 28)
    assert (p3.passes, p3.bits1, p3.bits2) == (2, 7, 6)
    p4 = pj.plan_default(1 << 22, 1 << 24)
    assert (p4.passes, p4.bits1, p4.bits2) == (2, 6, 5)
    p8 = pj.plan_default(1 << 26, 1 << 28, 8)
    assert (p8.passes, p8.bits1, p8.bits2) == (2, 8, 8)
    assert (pj.plan_default((1 << 26) + 1, 1 << 28).bits1, pj.plan_default((1 << 26) + 1, 1 << 28).bits2) == (8, 8)
    tiny = pj.plan_default(5, 5)
    assert tiny.passes == 1 and tiny.bits1 >= 1
    huge = pj.plan_default((1 << 31) - 1, 1 << 20)
    assert huge.bits1 + huge.bits2 == 16
    for n in [1, 100, 10**4, 10**6, 10**7, 10**8]:
        p = pj.plan_default(n, n, 8)
        assert 1 <= p.bits1 <= 12 and p.max_partition_size_factor == 1.25 and p.source == 0
        assert p.passes == (1 if p.bits2 == 0 else 2)


def test_workspace_size_monotone(pj):
    a = pj.workspace_size(1 << 20, 1 << 22)
    b = pj.workspace_size(1 << 22, 1 << 24)
    c = pj.workspace_size(1 << 22, 1 << 24, 8)
    assert 0 < a < b < c
    # two passes need the second tuple buffers
    one = pj.workspace_size(1 << 20, 1 << 20, 4, {"passes": 1, "bits1": 10})
    two = pj.workspace_size(1 << 20, 1 << 20, 4, {"passes": 2, "bits1": 5, "bits2": 5})
    assert two > one + 2 * (1 << 20) * 8


def _tables(lib_mod, nb=100, npr=100, kb=4, kb2=None):
    T = lib_mod.PjTable
    return T(1 << 12, None, nb, kb, 0), T(1 << 12, None, npr, kb if kb2 is None else kb2, 0)


@pytest.mark.parametrize("case,status", [
    ("bad_key_bytes", 1), ("negative_n", 1), ("mixed_widths", 1), ("bad_passes", 1), ("pass_bits_over_12", 1),
    ("bits2_one_pass", 1), ("small_workspace", 3), ("null_out", 1), ("bad_factor", 1)])
def test_argument_errors_without_device_work(pj, case, status):
    from partitioned_join_amd import _lib
    L = pj.lib()
    b, p = _tables(_lib)
    plan = None
    out, cap, ws, wsb = 1 << 12, 10, 1 << 12, 1 << 40
    if case == "bad_key_bytes":
        b.key_bytes = p.key_bytes = 3
    elif case == "negative_n":
        b.n = -1
    elif case == "mixed_widths":
        p.key_bytes = 8
    elif case == "bad_passes":
        plan = _lib.PjPlan(3, 4, 4, 0, 3.0, 0, 2)
    elif case == "pass_bits_over_12":
        plan = _lib.PjPlan(2, 13, 4, 0, 3.0, 0, 2)
    elif case == "bits2_one_pass":
        plan = _lib.PjPlan(1, 4, 4, 0, 3.0, 0, 2)
    elif case == "small_workspace":
        wsb = 1024
    elif case == "null_out":
        out = None
    elif case == "bad_factor":
        plan = _lib.PjPlan(1, 4, 0, 0, 0.0, 0, 2)
    cnt = ctypes.c_int64(0)
    st = L.pj_partitioned_join(ctypes.byref(b), ctypes.byref(p), ctypes.byref(plan) if plan else None, out, cap,
                               ctypes.byref(cnt), ws, wsb, None)
    assert st == status, L.pj_last_error()
    assert L.pj_last_error().decode()


def test_zipf_lut_matches_oracle(pj, oracle):
    from partitioned_join_amd import generators
    for n, theta in [(1, 1.0), (1000, 0.5), (4096, 1.0), (10000, 1.5)]:
        assert np.array_equal(generators.zipf_lut(n, theta), oracle.zipf_lut(n, theta))


def test_phase_workspace_sizes(pj):
    L = pj.lib()
    assert L.pj_histogram_workspace_size(16) > 0 and L.pj_histogram_workspace_size(17) == 0
    assert L.pj_scan_workspace_size(1 << 20) > 0
    assert L.pj_partition2_workspace_size(1000, 8, 8) > 0 and L.pj_partition2_workspace_size(1000, 13, 1) == 0
    assert L.pj_join_workspace_size(16) > L.pj_join_workspace_size(4) > 0
    assert L.pj_assign_blocks_workspace_size(65536) > 0
