"""GPU parity of the device primitives and of each partitioning phase, called
through the C ABI and checked against the oracle / golden vectors."""
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


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def u32(t):
    return t.cpu().numpy().view(np.uint32).astype(np.int64)


def test_hash_keys_match_oracle(pj, oracle, torch):
    # the oracle's MurmurHash3_x86_32 is pinned by SMHasher's verification value
    # (tests/test_oracle.py); edge keys plus seeded random ones of both widths
    rng = np.random.default_rng(0x4A5)
    k4 = np.concatenate([np.array([0, 1, -1, 2**31 - 1, -2**31, 0x1234567, -42], dtype=np.int32),
                         rng.integers(-2**31, 2**31 - 1, 2000, dtype=np.int64).astype(np.int32)])
    got4 = u32(pj.hash_keys(torch.from_numpy(k4).cuda()))
    assert got4.tolist() == [oracle.murmur3_key(int(k), 4) for k in k4]
    k8 = np.concatenate([np.array([0, 1, -1, 2**63 - 1, -2**63, 2**32, 0x123456789AB], dtype=np.int64),
                         rng.integers(-2**63, 2**63 - 1, 2000, dtype=np.int64)])
    got8 = u32(pj.hash_keys(torch.from_numpy(k8).cuda()))
    assert got8.tolist() == [oracle.murmur3_key(int(k), 8) for k in k8]


def test_fastdiv_matches_c_division(pj, torch):
    # C's truncating / and % (the semantics int_fastdiv reproduces,
    # src/int_fastdiv.h:130-133) over edge numerators and divisors
    edge = [0, 1, -1, 2, -2, 3, 7, 100, -100, 6143, 6144, 6145, 2**31 - 1, -2**31, -2**31 + 1, 2**30, -2**30]
    rng = np.random.default_rng(0xFD)
    nn = np.array(edge + rng.integers(-2**31, 2**31 - 1, 4000).tolist(), dtype=np.int64)
    n = torch.from_numpy(nn.astype(np.int32)).cuda()
    def c_div(x, d):
        q = abs(x) // abs(d)
        return q if (x >= 0) == (d > 0) else -q

    for d in [1, -1, 2, -2, 3, 7, 32, 6144, 5568, -6144, 1000003, 2**30, 2**31 - 1, -2**31 + 1]:
        q, r = pj.fastdiv(n, d)
        ok = ~((nn == -2**31) & (d == -1))  # overflows in C as well
        eq = np.array([c_div(int(x), d) if o else 0 for x, o in zip(nn, ok)], dtype=np.int64)
        er = nn - eq * d
        gq, gr = q.cpu().numpy().astype(np.int64), r.cpu().numpy().astype(np.int64)
        assert np.array_equal(gq[ok], eq[ok]), d
        assert np.array_equal(gr[ok], er[ok]), d


def test_generators_bit_exact(pj, oracle, torch):
    G = pj.generators
    assert np.array_equal(G.build_unique_i32(70001, 90000, 0x1D).cpu().numpy(),
                          oracle.gen_build_unique_i32(70001, 90000, 0x1D))
    assert np.array_equal(G.build_dup_i32(5000, 300, 1000, 3, 4).cpu().numpy(),
                          oracle.gen_build_dup_i32(5000, 300, 1000, 3, 4))
    assert np.array_equal(G.probe_i32(50000, 4096, 4096, 7, 8, 0.5, 4096, 4096).cpu().numpy(),
                          oracle.gen_probe_i32(50000, 4096, 4096, 7, 8, 0.5, 4096, 4096))
    assert np.array_equal(G.probe_i32(50000, 9000, 4096, 7, 8, 0.3).cpu().numpy(),
                          oracle.gen_probe_i32(50000, 9000, 4096, 7, 8, 0.3))
    lut = oracle.zipf_lut(10000, 1.0)
    assert np.array_equal(G.probe_zipf_i32(40000, 10000, 1.0, 5, 6).cpu().numpy(),
                          oracle.gen_probe_zipf_i32(40000, lut, 5, 6))
    assert np.array_equal(G.build_i64(3000, 11).cpu().numpy(), oracle.gen_build_i64(3000, 11))
    assert np.array_equal(G.probe_i64(9000, 3000, 11, 12, 0.4).cpu().numpy(), oracle.gen_probe_i64(9000, 3000, 11, 12, 0.4))
    assert np.array_equal(G.probe_zipf_i64(9000, 3000, 1.2, 11, 13, 14).cpu().numpy(),
                          oracle.gen_probe_zipf_i64(9000, oracle.zipf_lut(3000, 1.2), 11, 13, 14))


@pytest.mark.parametrize("n,total_bits,bits2", [(1, 4, 2), (1000, 8, 4), (300000, 10, 0), (2000000, 15, 7),
                                                (3000000, 16, 8), (777777, 16, 4)])
def test_combined_histogram(pj, oracle, torch, n, total_bits, bits2):
    rng = np.random.default_rng(n)
    keys = rng.integers(-2**31, 2**31 - 1, n).astype(np.int32)
    d = torch.from_numpy(keys).cuda()
    fine, p1 = pj.combined_histogram_pass(d, total_bits, bits2)
    assert np.array_equal(u32(fine), oracle.histogram(keys, 0, total_bits))
    assert np.array_equal(u32(p1), oracle.histogram(keys, bits2, total_bits - bits2))


def test_combined_histogram_skewed_16bit_overflow_epochs(pj, oracle, torch):
    # one bin receives far more than 65535 keys inside a single segment
    n = 3 << 20
    keys = np.full(n, 0x12345, dtype=np.int32)
    keys[::7] = np.arange(len(keys[::7]), dtype=np.int32)
    fine, p1 = pj.combined_histogram_pass(torch.from_numpy(keys).cuda(), 16, 8)
    assert np.array_equal(u32(fine), oracle.histogram(keys, 0, 16))
    assert np.array_equal(u32(p1), oracle.histogram(keys, 8, 8))


def test_combined_histogram_int64(pj, oracle, torch):
    rng = np.random.default_rng(5)
    keys = rng.integers(-2**63, 2**63 - 1, 500000, dtype=np.int64)
    fine, p1 = pj.combined_histogram_pass(torch.from_numpy(keys).cuda(), 12, 6)
    assert np.array_equal(u32(fine), oracle.histogram(keys, 0, 12))
    assert np.array_equal(u32(p1), oracle.histogram(keys, 6, 6))


@pytest.mark.parametrize("shift,bits", [(0, 8), (8, 8), (5, 11), (20, 12)])
def test_histogram_pass(pj, oracle, torch, shift, bits):
    rng = np.random.default_rng(bits)
    keys = rng.integers(-2**31, 2**31 - 1, 123457).astype(np.int32)
    h = pj.histogram_pass(torch.from_numpy(keys).cuda(), shift, bits)
    assert np.array_equal(u32(h), oracle.histogram(keys, shift, bits))


@pytest.mark.parametrize("n", [1, 5, 1023, 16384, 16385, 1000003])
def test_scans(pj, oracle, torch, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 1000, n).astype(np.int32)
    d = torch.from_numpy(x).cuda()
    assert np.array_equal(u32(pj.exclusive_scan(d)), oracle.exclusive_scan(x) % 2**32)
    assert np.array_equal(u32(pj.inclusive_scan(d)), oracle.inclusive_scan(x) % 2**32)


def tuples_np(t, key_bytes):
    a = t.cpu().numpy()
    if key_bytes == 4:
        return a[:, 0].copy(), a[:, 1].copy()
    keys = a[:, :2].copy().view(np.int64).reshape(-1)
    return keys, a[:, 2].copy()


@pytest.mark.parametrize("n,shift,bits,with_pay", [(0, 0, 4, False), (1, 0, 1, False), (100, 0, 3, True),
                                                   (65536 + 17, 0, 8, False), (1 << 20, 4, 10, True),
                                                   (3000001, 0, 12, False), (2000000, 9, 7, True),
                                                   (1500007, 0, 9, True)])
def test_partition_pass_one(pj, oracle, torch, n, shift, bits, with_pay):
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(-2**31, 2**31 - 1, n).astype(np.int32)
    pay = rng.integers(0, 2**31 - 1, n).astype(np.int32) if with_pay else None
    dk = torch.from_numpy(keys).cuda()
    dp = torch.from_numpy(pay).cuda() if with_pay else None
    tup, off = pj.partition_input_tables_pass_one(dk, dp, bits=bits, shift=shift)
    k, p = tuples_np(tup, 4)
    oracle.check_partitioned(keys, pay, k, p, off.cpu().numpy().astype(np.int64), shift, bits)
    assert int(off[-1]) == n


def test_partition_pass_one_int64(pj, oracle, torch):
    rng = np.random.default_rng(3)
    keys = rng.integers(-2**63, 2**63 - 1, 400000, dtype=np.int64)
    tup, off = pj.partition_input_tables_pass_one(torch.from_numpy(keys).cuda(), None, bits=9, shift=3)
    k, p = tuples_np(tup, 8)
    oracle.check_partitioned(keys, None, k, p, off.cpu().numpy().astype(np.int64), 3, 9)


def test_partition_pass_one_skewed(pj, oracle, torch):
    n = 1 << 21
    keys = np.full(n, 77, dtype=np.int32)
    keys[::3] = np.arange(len(keys[::3]), dtype=np.int32)
    tup, off = pj.partition_input_tables_pass_one(torch.from_numpy(keys).cuda(), None, bits=8)
    k, p = tuples_np(tup, 4)
    oracle.check_partitioned(keys, None, k, p, off.cpu().numpy().astype(np.int64), 0, 8)


# pass-2 fan-outs 4..4096 cover every scatter tile / line-carry configuration
@pytest.mark.parametrize("n,b1,b2,kb", [(1000, 3, 2, 4), (1 << 20, 8, 8, 4), (777777, 6, 9, 4), (300000, 5, 5, 8),
                                        (900001, 4, 10, 4), (1 << 20, 3, 12, 4), (500009, 2, 11, 8)])
def test_partition_pass_two(pj, oracle, torch, n, b1, b2, kb):
    rng = np.random.default_rng(n)
    dt = np.int32 if kb == 4 else np.int64
    keys = rng.integers(np.iinfo(dt).min, np.iinfo(dt).max, n, dtype=dt)
    t1, off1 = pj.partition_input_tables_pass_one(torch.from_numpy(keys).cuda(), None, bits=b1, shift=b2)
    t2, off2 = pj.partition_input_tables_pass_two(t1, kb, b1, b2, off1)
    k, p = tuples_np(t2, kb)
    oracle.check_partitioned(keys, None, k, p, off2.cpu().numpy().astype(np.int64), 0, b1 + b2)


def test_assign_blocks(pj, oracle, torch):
    from tests.test_oracle import emulate_reference_block_assignment
    rng = np.random.default_rng(12)
    for P, base in [(1, 5), (17, 3), (4096, 100), (65536, 3072)]:
        sizes = rng.integers(0, 2 * base, P).astype(np.int64)
        sizes[rng.integers(0, P, max(1, P // 100))] *= 20  # heavy partitions
        offsets = oracle.exclusive_scan(sizes)
        total = int(sizes.sum())
        scanned, pidx, tot = pj.handle_skewed_data(torch.from_numpy(offsets.astype(np.int32)).cuda(), total, P, base)
        e_scanned, e_pidx, e_nblk = emulate_reference_block_assignment(offsets.tolist(), total, base)
        assert int(tot.item()) == e_nblk
        assert u32(scanned).tolist() == e_scanned
        assert u32(pidx)[:e_nblk].tolist() == e_pidx


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_join_of_partitions_not_split_by_low_bits(pj, oracle, torch, dtype):
    """pj_join_partitions over tuples partitioned on bits 4..7 (shift 4) and
    joined as 16 partitions: partition q then holds keys whose low bits are
    not q, including the value q ^ 1 (key 1 lies in partition 0).  Every key
    value must still join, duplicates included."""
    rng = np.random.default_rng(31)
    bk = rng.integers(-40, 300, 20000).astype(dtype)
    pk = rng.integers(-40, 300, 70000).astype(dtype)
    bk[:6] = [0, 1, 2, 3, -1, 17]
    pk[:6] = [1, 0, 3, 2, -1, 16]
    d = lambda a: torch.from_numpy(a).cuda()
    kb = bk.itemsize
    tb, ob = pj.partition_input_tables_pass_one(d(bk), None, bits=4, shift=4)
    tp, op = pj.partition_input_tables_pass_one(d(pk), None, bits=4, shift=4)
    exp = oracle.sort_pairs(oracle.inner_join_cpu(bk, pk, uniq_keys=False))
    for base in (1000, 2783 if kb == 4 else 2047):
        pairs, n = pj.shared_memory_hash_join(tb, ob, tp, op, kb, 4, base, out_capacity=len(exp) + 64)
        assert n == len(exp), (base, n, len(exp))
        assert np.array_equal(oracle.sort_pairs(pairs.cpu().numpy()), exp), base


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_join_of_a_chunk_holding_both_extreme_keys(pj, oracle, torch, dtype):
    """One caller-made partition whose build chunk holds the smallest and the
    largest key value (as unsigned: 0 and all ones) and 1, 2, 3, ...: no value
    below or above the chunk is free, so the empty marker is the smallest
    missing value of 1..c+1."""
    info = np.iinfo(dtype)
    nb = 1500
    bk = np.concatenate([np.array([0, -1, info.min, info.max], dtype=dtype), np.arange(1, nb - 3, dtype=dtype)])
    bk = np.delete(bk, 700)  # leave a gap at 697 (the marker must be a value absent from the chunk)
    pk = np.concatenate([bk, np.array([697, 5000, 0, -1, info.max], dtype=dtype)])
    rng = np.random.default_rng(5)
    bk, pk = rng.permutation(bk), rng.permutation(pk)
    kb = bk.itemsize
    cols = 2 if kb == 4 else 4

    def tuples(keys):
        t = np.zeros((len(keys), cols), dtype=np.int32)
        if kb == 4:
            t[:, 0] = keys
        else:
            t[:, 0:2] = keys.astype(np.int64).view(np.int32).reshape(-1, 2)
        t[:, cols - 2 if kb == 8 else 1] = np.arange(len(keys), dtype=np.int32)
        return torch.from_numpy(t).cuda()

    # total_bits 1: partition 0 holds every tuple, partition 1 none
    ob = torch.tensor([0, len(bk), len(bk)], dtype=torch.int32, device="cuda")
    op = torch.tensor([0, len(pk), len(pk)], dtype=torch.int32, device="cuda")
    exp = oracle.sort_pairs(oracle.inner_join_cpu(bk, pk, uniq_keys=False))
    pairs, n = pj.shared_memory_hash_join(tuples(bk), ob, tuples(pk), op, kb, 1, 2047, out_capacity=len(exp) + 8)
    assert n == len(exp)
    assert np.array_equal(oracle.sort_pairs(pairs.cpu().numpy()), exp)


def test_shared_memory_hash_join_phase(pj, oracle, torch):
    rng = np.random.default_rng(13)
    nb, npr, B = 200000, 600000, 10
    bk = rng.permutation(1 << 20)[:nb].astype(np.int32)
    pk = rng.integers(0, 1 << 20, npr).astype(np.int32)
    tb, ob = pj.partition_input_tables_pass_one(torch.from_numpy(bk).cuda(), None, bits=B)
    tp, op = pj.partition_input_tables_pass_one(torch.from_numpy(pk).cuda(), None, bits=B)
    base = oracle.base_build_size(nb, 1 << B)
    pairs, n = pj.shared_memory_hash_join(tb, ob, tp, op, 4, B, base)
    exp = oracle.inner_join_cpu(bk, pk)
    assert n == len(exp)
    assert np.array_equal(oracle.sort_pairs(pairs.cpu().numpy()), oracle.sort_pairs(exp))
