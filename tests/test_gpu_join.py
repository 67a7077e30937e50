"""GPU parity of the full join (pj_partitioned_join through the C ABI) against
the oracle's CPU join: sorted (build_idx, probe_idx) multisets must be equal.
Covers the benchmark configurations at oracle-checkable sizes, the reference's
own edge cases and, at full size, size-independent properties."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def sorted_np(p):
    p = np.asarray(p, dtype=np.int64).reshape(-1, 2)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def check_join(pj, oracle, torch, bk, pk, bp=None, pp=None, plan=None, uniq=True):
    d = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda()
    got = pj.partitioned_join(d(bk), d(pk), d(bp), d(pp), plan=plan)
    exp = oracle.inner_join_cpu(bk, pk, bp, pp, uniq_keys=uniq)
    g = oracle.sort_pairs(got.cpu().numpy())
    e = oracle.sort_pairs(exp)
    assert len(g) == len(e), (len(g), len(e))
    assert np.array_equal(g, e)
    return g


def test_known_answers(pj, oracle, torch):
    with open(os.path.join(GOLDEN, "cudf_kat.json")) as f:
        cases = json.load(f)
    for c in cases:
        bk = np.array(c["build"], dtype=np.int32)
        pk = np.array(c["probe"], dtype=np.int32)
        d = lambda a: torch.from_numpy(a).cuda()
        got = pj.partitioned_join(d(bk), d(pk))
        assert np.array_equal(sorted_np(got.cpu().numpy()), sorted_np(c["pairs"])), c["name"]


def test_reference_output_projection(pj, oracle, torch):
    """The reference's output form {key, probe payload} (src/hash_join_kernels.cuh:174-176)
    is the projection (R.key[b], S.payload[p]) of the pairs: checked against the
    oracle's projection of its own join, with duplicates on both sides and the
    reference's zero probe payload (main.cu's driver) as well as a payload column."""
    rng = np.random.default_rng(0x9E0)
    for nb, ns, span, uniq in [(50, 200, 40, 0), (1000, 5000, 3000, 0), (4096, 4096, 1 << 30, 1)]:
        bk = rng.integers(-span, span, nb).astype(np.int32)
        if uniq:
            bk = rng.permutation(np.arange(nb, dtype=np.int32) * 7 - 11)
        pk = rng.integers(-span, span, ns).astype(np.int32)
        pp = rng.integers(-2**31, 2**31 - 1, ns).astype(np.int32)
        for pay in (None, pp):
            dbk, dpk = torch.from_numpy(bk).cuda(), torch.from_numpy(pk).cuda()
            dpp = None if pay is None else torch.from_numpy(pay).cuda()
            pairs = pj.partitioned_join(dbk, dpk, None, dpp, plan={"passes": 1, "bits1": 4, "unique_build_keys": uniq})
            exp = oracle.inner_join_cpu(bk, pk, None, pay, uniq_keys=bool(uniq))
            assert pairs.shape[0] == len(exp)
            # both sides' pairs hold build row ids and probe payloads (row ids without a column)
            xy = pj.project_reference(dbk, pairs).cpu().numpy()
            ref = oracle.project_to_reference(exp, bk)
            assert np.array_equal(sorted_np(xy), sorted_np(ref))


def test_config1_exact(pj, oracle, torch):
    """1M unique build keys (perm of [0, 2^20)) x 4M probes drawn from them:
    exactly 4,194,304 pairs, every probe row exactly once."""
    bk, pk = pj.generators.workload_config1()
    pairs = pj.partitioned_join(bk, pk)
    assert pairs.shape[0] == 1 << 22
    b, p = bk.cpu().numpy(), pk.cpu().numpy()
    exp = oracle.inner_join_cpu(b, p)
    assert np.array_equal(oracle.sort_pairs(pairs.cpu().numpy()), oracle.sort_pairs(exp))
    assert np.array_equal(np.sort(pairs[:, 1].cpu().numpy()), np.arange(1 << 22))


@pytest.mark.parametrize("plan", [None, {"passes": 1, "bits1": 6}, {"passes": 2, "bits1": 5, "bits2": 7},
                                  {"passes": 2, "bits1": 8, "bits2": 8, "unique_build_keys": 1},
                                  {"passes": 1, "bits1": 12, "max_partition_size_factor": 1.0},
                                  {"passes": 2, "bits1": 3, "bits2": 3, "probe_slice": 1000}])
def test_config2_shape_scaled(pj, oracle, torch, plan):
    """Config-2 distribution at 1/16 size: 4M unique keys x 16M probes, 50 % match."""
    bk, pk = pj.generators.workload_config2(scale_shift=4)
    check_join(pj, oracle, torch, bk.cpu().numpy(), pk.cpu().numpy(), plan=plan)


@pytest.mark.parametrize("theta", [0.5, 1.0, 1.5])
def test_config3_zipf_scaled(pj, oracle, torch, theta):
    """Config-3 distribution at 1/16 size: 1M unique keys x 16M Zipf probes."""
    bk, pk = pj.generators.workload_config3(theta, scale_shift=4)
    g = check_join(pj, oracle, torch, bk.cpu().numpy(), pk.cpu().numpy())
    assert len(g) == pk.numel()


def test_duplicate_build_keys(pj, oracle, torch):
    rng = np.random.default_rng(21)
    bk = rng.integers(0, 5000, 200000).astype(np.int32)
    pk = rng.integers(0, 6000, 300000).astype(np.int32)
    check_join(pj, oracle, torch, bk, pk, uniq=False)
    check_join(pj, oracle, torch, bk, pk, plan={"passes": 2, "bits1": 4, "bits2": 4}, uniq=False)


def test_all_equal_keys_large_output(pj, oracle, torch):
    # like cuDF's HashJoinLargeOutputSize, scaled: 3000 x 4000 identical keys -> 12M pairs
    bk = np.zeros(3000, dtype=np.int32)
    pk = np.zeros(4000, dtype=np.int32)
    got = pj.partitioned_join(torch.from_numpy(bk).cuda(), torch.from_numpy(pk).cuda())
    assert got.shape[0] == 12_000_000
    g = got.cpu().numpy().astype(np.int64)
    code = np.sort(g[:, 0] * 4000 + g[:, 1])
    assert np.array_equal(code, np.arange(12_000_000))


def test_payload_columns(pj, oracle, torch):
    rng = np.random.default_rng(22)
    bk = rng.permutation(1 << 18).astype(np.int32)
    pk = rng.integers(0, 1 << 19, 1 << 20).astype(np.int32)
    bp = rng.integers(-2**31, 2**31 - 1, len(bk)).astype(np.int32)
    pp = rng.integers(-2**31, 2**31 - 1, len(pk)).astype(np.int32)
    check_join(pj, oracle, torch, bk, pk, bp, pp)
    check_join(pj, oracle, torch, bk, pk, bp, None)


def test_edge_keys_and_sizes(pj, oracle, torch):
    extremes = np.array([-1, 0, 1, -2**31, 2**31 - 1, -2, 65535, 65536, -65536], dtype=np.int32)
    bk = np.concatenate([extremes, np.arange(1000, dtype=np.int32) * 65536])
    pk = np.concatenate([extremes, extremes, np.arange(2000, dtype=np.int32) * 32768])
    # 0 and 65536 occur twice on the build side: non-unique oracle semantics
    for plan in [None, {"passes": 1, "bits1": 1}, {"passes": 2, "bits1": 8, "bits2": 8}]:
        check_join(pj, oracle, torch, bk, pk, plan=plan, uniq=False)
    one = np.array([7], dtype=np.int32)
    check_join(pj, oracle, torch, one, one)
    check_join(pj, oracle, torch, one, np.array([8], dtype=np.int32))


def test_empty_tables(pj, torch):
    e = torch.empty(0, dtype=torch.int32, device="cuda")
    x = torch.arange(10, dtype=torch.int32, device="cuda")
    assert pj.partitioned_join(e, x).shape[0] == 0
    assert pj.partitioned_join(x, e).shape[0] == 0
    assert pj.partitioned_join(e, e).shape[0] == 0


def test_int64_keys(pj, oracle, torch):
    G = pj.generators
    bk = G.build_i64(300000, 0x64)
    pk = G.probe_i64(900000, 300000, 0x64, 0x65, 0.6)
    check_join(pj, oracle, torch, bk.cpu().numpy(), pk.cpu().numpy())
    lut = oracle.zipf_lut(300000, 1.2)
    pz = G.probe_zipf_i64(900000, 300000, 1.2, 0x64, 0x66, 0x67, lut=lut)
    check_join(pj, oracle, torch, bk.cpu().numpy(), pz.cpu().numpy(), plan={"passes": 2, "bits1": 4, "bits2": 4})
    rng = np.random.default_rng(23)
    dup = rng.integers(-3, 3, 2000).astype(np.int64) * (2**40 + 1)
    check_join(pj, oracle, torch, dup, dup[::-1].copy(), uniq=False)


@pytest.mark.parametrize("plan", [
    {"passes": 2, "bits1": 6, "bits2": 6},   # pass 1: 8K tile + 128-digit carry; pass 2: 8K + 64-digit carry
    {"passes": 2, "bits1": 7, "bits2": 7},   # pass 2: 7K tile + 256-digit carry
    {"passes": 2, "bits1": 8, "bits2": 8},   # pass 1: 4K tile + 512-digit carry
    {"passes": 1, "bits1": 9},               # one pass, fan-out 512: 4K + 512-digit carry
    {"passes": 2, "bits1": 5, "bits2": 9},   # pass 2 fan-out 512: 4K + 512-digit carry
    {"passes": 2, "bits1": 4, "bits2": 10},  # pass 2 fan-out 1024: 8K, no carry
])
def test_int64_scatter_configurations(pj, oracle, torch, plan):
    """Every 16-byte-tuple scatter configuration (tile and line carry chosen
    by the fan-out of each pass) on 64-bit keys, against the oracle."""
    G = pj.generators
    bk = G.build_i64(300000, 0x65)
    pk = G.probe_i64(900000, 300000, 0x65, 0x66, 0.6)
    check_join(pj, oracle, torch, bk.cpu().numpy(), pk.cpu().numpy(), plan=plan)


@pytest.mark.parametrize("plan", [{"passes": 2, "bits1": 7, "bits2": 7}, {"passes": 1, "bits1": 9}])
def test_int64_payload_columns_and_duplicates(pj, oracle, torch, plan):
    """64-bit keys with duplicate build keys and int32 payload columns on both
    sides (the payload variants of the 16-byte-tuple scatters)."""
    rng = np.random.default_rng(24)
    bk = rng.integers(-2**62, 2**62, 150000).astype(np.int64)
    bk[1::2] = bk[0::2]  # every key twice
    bk[::500] = bk[3]    # and one key 300 times
    pk = np.concatenate([rng.choice(bk, 400000), rng.integers(-2**62, 2**62, 100000)]).astype(np.int64)
    bp = rng.integers(-2**31, 2**31 - 1, len(bk)).astype(np.int32)
    pp = rng.integers(-2**31, 2**31 - 1, len(pk)).astype(np.int32)
    check_join(pj, oracle, torch, bk, pk, bp, pp, plan=plan, uniq=False)


@pytest.mark.parametrize("plan", [{"passes": 2, "bits1": 9, "bits2": 9},
                                  {"passes": 2, "bits1": 12, "bits2": 5, "unique_build_keys": 1},
                                  {"passes": 2, "bits1": 6, "bits2": 12},
                                  {"passes": 2, "bits1": 12, "bits2": 12, "max_partition_size_factor": 1.0},
                                  {"passes": 2, "bits1": 12, "bits2": 7},
                                  {"passes": 2, "bits1": 12, "bits2": 8, "unique_build_keys": 1},
                                  {"passes": 2, "bits1": 12, "bits2": 10}])
def test_deep_plans_beyond_the_combined_histogram(pj, oracle, torch, plan):
    """More than 2^16 partitions (the reference's skew script uses 512 x 512):
    pass-2 digits are counted per item after pass 1.  Unique and duplicate
    build keys, both key widths, against the oracle."""
    rng = np.random.default_rng(plan["bits1"] * 100 + plan["bits2"])
    for dt in (np.int32, np.int64):
        bk = rng.permutation(1 << 20).astype(dt) if plan.get("unique_build_keys") else \
            rng.integers(0, 1 << 19, 1 << 20).astype(dt)
        pk = rng.integers(0, 1 << 21, 3 << 20).astype(dt)
        if dt == np.int64:
            bk = bk * (2**32 + 3)
            pk = pk * (2**32 + 3)
        check_join(pj, oracle, torch, bk, pk, plan=plan, uniq=bool(plan.get("unique_build_keys")))


def test_capacity_error_reports_count(pj, torch):
    from partitioned_join_amd import PjError
    bk = torch.arange(1000, dtype=torch.int32, device="cuda")
    pk = torch.arange(5000, dtype=torch.int32, device="cuda") % 1000
    with pytest.raises(PjError) as ei:
        pj.partitioned_join(bk, pk, out_capacity=100)
    assert ei.value.status == 2 and ei.value.count == 5000
    ok = pj.partitioned_join(bk, pk, out_capacity=5000)
    assert ok.shape[0] == 5000


def test_repeat_calls_reuse_workspace(pj, oracle, torch):
    bk, pk = pj.generators.workload_config2(scale_shift=6)
    a = oracle.sort_pairs(pj.partitioned_join(bk, pk).cpu().numpy())
    b = oracle.sort_pairs(pj.partitioned_join(bk, pk).cpu().numpy())
    small_b, small_p = bk[:1000].clone(), pk[:5000].clone()
    c = oracle.sort_pairs(pj.partitioned_join(small_b, small_p).cpu().numpy())
    assert np.array_equal(a, b)
    assert np.array_equal(c, oracle.sort_pairs(oracle.inner_join_cpu(small_b.cpu().numpy(), small_p.cpu().numpy())))


@pytest.mark.parametrize("unique", [0, 1])
def test_config2_full_size_properties(pj, torch, unique):
    """64M x 256M at full size, checked on the device through properties that
    do not need the CPU oracle: the pair count equals the number of probe keys
    below 2^26, every such probe row appears exactly once, and every pair joins
    equal keys.  With and without the plan's unique-key promise (the bench's
    plan declares it)."""
    bk, pk = pj.generators.workload_config2()
    print(f"64M x 256M, unique_build_keys={unique}", flush=True)
    plan = pj.plan_default(bk.numel(), pk.numel())  # the bench's plan: 8 + 7 bits, chunk factor 1.25
    plan.unique_build_keys = unique
    pairs, n = pj.partitioned_join(bk, pk, plan=plan, return_count=True)
    match = pk < (1 << 26)
    assert n == int(match.sum().item())
    b = pairs[:, 0].long()
    p = pairs[:, 1].long()
    assert bool((bk[b] == pk[p]).all().item())
    seen = torch.zeros(pk.numel(), dtype=torch.int32, device="cuda")
    seen.index_add_(0, p, torch.ones_like(pairs[:, 1]))
    assert bool((seen == match.int()).all().item())


@pytest.mark.parametrize("theta", [0.5, 1.0, 1.5])
def test_config3_full_size_properties(pj, torch, theta):
    """16M x 256M Zipf(theta) at full size, checked on the device: every probe
    key is a build key, so the count is exactly 2^28, every probe row appears
    exactly once, and every pair joins equal keys (heavy keys split across many
    work items on both sides)."""
    bk, pk = pj.generators.workload_config3(theta)
    print(f"theta {theta}: 16M x 256M", flush=True)
    pairs, n = pj.partitioned_join(bk, pk, return_count=True)
    assert n == 1 << 28
    b = pairs[:, 0].long()
    p = pairs[:, 1].long()
    assert bool((bk[b] == pk[p]).all().item())
    seen = torch.zeros(pk.numel(), dtype=torch.int32, device="cuda")
    seen.index_add_(0, p, torch.ones_like(pairs[:, 1]))
    assert bool((seen == 1).all().item())
    del pairs, b, p, seen


def test_output_beyond_2_to_32_pairs_reports_exact_count(pj, torch):
    """65567 equal keys on both sides join to 65567^2 = 4,299,031,489 pairs,
    beyond 2^32 (cuDF HashJoinLargeOutputSize, cudf:tests/join/join_tests.cpp):
    with a small output the call returns PJ_ERR_CAPACITY and the exact 64-bit
    count, for 4- and 8-byte keys."""
    n = 65567
    for dt in (torch.int32, torch.int64):
        bk = torch.full((n,), 7, dtype=dt, device="cuda")
        pk = torch.full((n,), 7, dtype=dt, device="cuda")
        with pytest.raises(pj.PjError) as e:
            pj.partitioned_join(bk, pk, out_capacity=1 << 20)
        assert e.value.status == 2 and e.value.count == n * n, (dt, e.value.count)


def test_gather_columns(pj, oracle, torch):
    """Gather map: payload columns of the joined rows, int32/int64/float64."""
    rng = np.random.default_rng(24)
    bk = rng.permutation(1 << 16).astype(np.int32)
    pk = rng.integers(0, 1 << 17, 300000).astype(np.int32)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    pairs = pj.partitioned_join(d(bk), d(pk))
    pr = pairs.cpu().numpy().astype(np.int64)
    bcol = rng.integers(-2**40, 2**40, len(bk)).astype(np.int64)
    pcol = rng.integers(-2**31, 2**31 - 1, len(pk)).astype(np.int32)
    gb, gp = pj.gather(pairs, d(bcol), d(pcol))
    assert np.array_equal(gb.cpu().numpy(), bcol[pr[:, 0]])
    assert np.array_equal(gp.cpu().numpy(), pcol[pr[:, 1]])
    fcol = rng.standard_normal(len(pk))
    none_b, gf = pj.gather(pairs, None, d(fcol))
    assert none_b is None and np.array_equal(gf.cpu().numpy(), fcol[pr[:, 1]])
    # every gathered build key equals the probe key it joined
    kb, kp = pj.gather(pairs, d(bk), d(pk))
    assert bool((kb == kp).all().item())
    e0, e1 = pj.gather(pairs[:0], d(bcol), d(pcol))
    assert e0.numel() == 0 and e1.numel() == 0


def test_profile_reads_joins_queued_back_to_back(pj, torch):
    """Phase times of joins queued without a host wait: one set per joined
    pair of tables (an empty table records none), oldest first, no phase
    negative and the phases summing to the total."""
    bk, pk = pj.generators.workload_config2(scale_shift=6)
    out = torch.empty((pk.numel() + 1024, 2), dtype=torch.int32, device="cuda")
    ws = torch.empty(pj.workspace_size(bk.numel(), pk.numel(), 4), dtype=torch.uint8, device="cuda")
    empty = torch.empty(0, dtype=torch.int32, device="cuda")
    pj.profile_enable(True)
    try:
        for k in range(3):
            pj.partitioned_join_async(bk, pk, out, workspace=ws)
            if k == 1:
                pj.partitioned_join_async(empty, pk, out, workspace=ws)
        pj.partitioned_join_async(bk, pk, out, workspace=ws)
        rec = pj.profile_read_recent(10)
        assert len(rec) == 4
        for r in rec:
            phases = [r[k] for k in ("histogram_ms", "offsets_ms", "pass1_ms", "pass2_ms", "skew_ms", "join_ms",
                                     "compact_ms")]
            assert r["valid"] == 1 and all(v >= 0 for v in phases) and r["join_ms"] > 0
            assert abs(sum(phases) - r["total_ms"]) <= 1e-3 * r["total_ms"] + 1e-3
        last = pj.profile_read()
        assert last["valid"] == 1 and last["total_ms"] == rec[-1]["total_ms"]
        assert len(pj.profile_read_recent(2)) == 2
        # only the two events around the join phase: the other phases read 0
        pj.profile_select(0x60)
        pj.partitioned_join_async(bk, pk, out, workspace=ws)
        sel = pj.profile_read_recent(1)[0]
        assert sel["valid"] == 1 and sel["join_ms"] > 0
        assert sel["histogram_ms"] == 0 and sel["pass1_ms"] == 0 and sel["total_ms"] == 0
        assert len(pj.profile_read_recent(10)) == 5  # the earlier sets keep their own events
    finally:
        pj.profile_select(0xFF)
        pj.profile_enable(False)
    assert pj.profile_read_recent(4) == []
