"""Pins of the CPU oracle (runs without a GPU).

Each oracle function is checked against something independent of its own
formula: the SMHasher verification value, C semantics of / and %, brute force
on tiny inputs, closed forms and cuDF's inner-join known answers (golden
vectors in tests/golden/).  The reference's own CPU code is not built in this
image (its headers need the CUDA toolkit), so nothing here runs it.
"""
import json
import math
import os
import struct

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


# ----------------------------------------------------------------- MurmurHash3
def test_murmur3_smhasher_verification(oracle):
    # SMHasher VerificationTest: hash keys 0..i-1 (i < 256) with seed 256-i,
    # hash the concatenated 32-bit results with seed 0.
    key = bytes(range(256))
    hashes = b"".join(struct.pack("<I", oracle.murmur3_32(key[:i], 256 - i)) for i in range(256))
    assert oracle.murmur3_32(hashes, 0) == 0xB0F57EE3


def test_murmur3_key_forms_equal_byte_form(oracle):
    rng = np.random.default_rng(1)
    for k in rng.integers(-2**31, 2**31 - 1, 200).tolist():
        assert oracle.murmur3_key(k, 4) == oracle.murmur3_32(struct.pack("<i", k), 0)
    for k in rng.integers(-2**62, 2**62, 200).tolist():
        assert oracle.murmur3_key(k, 8) == oracle.murmur3_32(struct.pack("<q", k), 0)


# ----------------------------------------------------------------- int_fastdiv
def c_div(n, d):
    q = abs(n) // abs(d)
    return q if (n >= 0) == (d > 0) else -q


def test_fastdiv_matches_c_semantics(oracle):
    rng = np.random.default_rng(3)
    divisors = [1, -1, 2, -2, 3, -3, 5, 7, 641, 6144, 2**31 - 1, -(2**31 - 1), -2**31 + 1] + \
        rng.integers(2, 2**31 - 1, 40).tolist() + (-rng.integers(2, 2**31 - 1, 10)).tolist()
    nums = [0, 1, -1, 2**31 - 1, -2**31, -2**31 + 1] + rng.integers(-2**31, 2**31 - 1, 300).tolist()
    for d in divisors:
        f = oracle.fastdiv(d)
        for n in nums:
            if n == -2**31 and d == -1:
                continue  # overflows in C as well
            q = oracle.fastdiv_div(n, f)
            assert q == c_div(n, d), (n, d)
            assert oracle.fastdiv_mod(n, f) == n - q * d


def test_fastdiv_small_exhaustive(oracle):
    for d in list(range(2, 70)) + list(range(-69, -1)):
        f = oracle.fastdiv(d)
        for n in range(-3000, 3000):
            assert oracle.fastdiv_div(n, f) == c_div(n, d)


def test_hash_slot_closed_form(oracle):
    rng = np.random.default_rng(4)
    for size in [2, 3, 2048, 6144, 6145, 7936]:
        for h in rng.integers(0, 2**32, 200).tolist():
            signed = h - 2**32 if h >= 2**31 else h
            assert oracle.hash_slot(h, size) == abs(signed - c_div(signed, size) * size)
            assert 0 <= oracle.hash_slot(h, size) < size


# ----------------------------------------------------------------- joins
def sorted_pairs(p):
    p = np.asarray(p, dtype=np.int64).reshape(-1, 2)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def test_join_known_answers(oracle):
    for case in load("cudf_kat.json"):
        b = np.array(case["build"], dtype=np.int32)
        p = np.array(case["probe"], dtype=np.int32)
        exp = sorted_pairs(case["pairs"])
        assert np.array_equal(sorted_pairs(oracle.inner_join_cpu(b, p, uniq_keys=False)), exp), case["name"]
        assert np.array_equal(sorted_pairs(oracle.sort_merge_join(b, p)), exp), case["name"]
        assert np.array_equal(sorted_pairs(oracle.nested_loop_join(b, p)), exp), case["name"]


@pytest.mark.parametrize("seed", range(6))
def test_join_vs_nested_loop(oracle, seed):
    rng = np.random.default_rng(seed)
    nb, npr = rng.integers(1, 60), rng.integers(1, 80)
    b = rng.integers(-5, 20, nb).astype(np.int32)
    p = rng.integers(-5, 20, npr).astype(np.int32)
    exp = sorted_pairs(oracle.nested_loop_join(b, p))
    assert np.array_equal(sorted_pairs(oracle.inner_join_cpu(b, p, uniq_keys=False)), exp)
    assert np.array_equal(sorted_pairs(oracle.sort_merge_join(b, p)), exp)
    assert oracle.join_count(b, p) == len(exp)
    bp = rng.integers(0, 1000, nb).astype(np.int32)
    pp = rng.integers(0, 1000, npr).astype(np.int32)
    exp2 = sorted_pairs(oracle.nested_loop_join(b, p, bp, pp))
    assert np.array_equal(sorted_pairs(oracle.inner_join_cpu(b, p, bp, pp, uniq_keys=False)), exp2)


def test_join_int64_keys(oracle):
    rng = np.random.default_rng(9)
    b = rng.integers(-2**62, 2**62, 50).astype(np.int64)
    p = np.concatenate([b[rng.integers(0, 50, 70)], rng.integers(-2**62, 2**62, 30)]).astype(np.int64)
    exp = sorted_pairs(oracle.nested_loop_join(b, p))
    assert np.array_equal(sorted_pairs(oracle.inner_join_cpu(b, p)), exp)
    assert np.array_equal(sorted_pairs(oracle.sort_merge_join(b, p)), exp)


def test_join_count_closed_form(oracle):
    rng = np.random.default_rng(5)
    b = rng.integers(0, 50, 3000).astype(np.int32)
    p = rng.integers(0, 60, 5000).astype(np.int32)
    cb = np.bincount(b, minlength=60)
    cp = np.bincount(p, minlength=60)
    assert oracle.join_count(b, p) == int((cb[:60] * cp[:60]).sum())


def _reference_form_by_definition(b, p):
    """{(R.key[b], S.payload[p]) : R.key[b] = S.key[p]} by nested loops: the
    reference's output form {x = key, y = probe payload}
    (src/hash_join_kernels.cuh:155-164), with row-id payloads."""
    return sorted_pairs([(int(kb), j) for kb in b.tolist() for j, kp in enumerate(p.tolist()) if kb == kp])


@pytest.mark.parametrize("uniq", [True, False])
def test_projection_equals_reference_form(oracle, uniq):
    rng = np.random.default_rng(13 if uniq else 14)
    b = (rng.permutation(300)[:200] if uniq else rng.integers(0, 150, 200)).astype(np.int32)
    p = rng.integers(0, 450, 700).astype(np.int32)
    pairs = oracle.inner_join_cpu(b, p, uniq_keys=uniq)
    assert np.array_equal(sorted_pairs(oracle.project_to_reference(pairs, b)), _reference_form_by_definition(b, p))


def test_project_to_reference_planted_mistake_is_caught(oracle):
    rng = np.random.default_rng(15)
    b = rng.permutation(300)[:200].astype(np.int32)
    p = rng.integers(0, 450, 700).astype(np.int32)
    pairs = oracle.inner_join_cpu(b, p, uniq_keys=True).copy()
    pairs[0, 0] = (pairs[0, 0] + 1) % len(b)  # planted: wrong build row
    assert not np.array_equal(sorted_pairs(oracle.project_to_reference(pairs, b)), _reference_form_by_definition(b, p))


# ----------------------------------------------------------------- generators
@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 1000, 4096, 5000, 65537])
def test_perm_is_bijection(oracle, n):
    v = oracle.gen_build_unique_i32(n, n, 0xABC)
    assert np.array_equal(np.sort(v), np.arange(n))


def test_build_unique_subset_of_domain(oracle):
    v = oracle.gen_build_unique_i32(3000, 10000, 7)
    assert len(np.unique(v)) == 3000 and v.min() >= 0 and v.max() < 10000


def test_probe_generator_semantics(oracle):
    nb, npr = 1 << 12, 1 << 16
    b = oracle.gen_build_unique_i32(nb, nb, 0xC2B)
    p = oracle.gen_probe_i32(npr, nb, nb, 0xC2B, 0xC2F, 0.5, nb, nb)
    inb = p < nb
    assert abs(inb.mean() - 0.5) < 0.01
    assert p.min() >= 0 and p.max() < 2 * nb
    assert oracle.join_count(b, p) == int(inb.sum())
    lot = oracle.gen_probe_i32(npr, 3 * nb, nb, 0x55, 0x66, 0.25)  # lottery mode
    b2 = oracle.gen_build_unique_i32(nb, 3 * nb, 0x55)
    assert abs(np.isin(lot, b2).mean() - 0.25) < 0.01
    assert lot.max() < 3 * nb


def test_zipf_lut_closed_form(oracle):
    n, theta = 1000, 1.0
    lut = oracle.zipf_lut(n, theta)
    h = sum(1.0 / i for i in range(1, n + 1))
    assert abs(lut[0] - 1.0 / h) < 1e-15
    assert abs(lut[-1] - 1.0) < 1e-12
    assert np.all(np.diff(lut) > 0)
    lut0 = oracle.zipf_lut(100, 0.0)
    assert np.allclose(lut0, np.arange(1, 101) / 100.0)


def test_zipf_rank_brute_force(oracle):
    lut = oracle.zipf_lut(50, 1.3)
    rng = np.random.default_rng(6)
    for u in list(rng.random(500)) + [0.0, lut[0], lut[10], 1.0 - 2**-53]:
        expect = next((i for i in range(50) if lut[i] >= u), 49)
        assert oracle.zipf_rank(lut, u) == expect


def test_zipf_heavy_mass(oracle):
    # top-1 mass at theta = 1 over 2^24 ranks is 1/H(2^24) ~ 5.8 %
    n = 1 << 24
    h = math.log(n) + 0.5772156649 + 1 / (2 * n)
    lut = oracle.zipf_lut(n, 1.0)
    assert abs(lut[0] - 1 / h) < 1e-6
    assert abs(lut[0] - 0.0581) < 0.001
    keys = oracle.gen_probe_zipf_i32(1 << 18, lut, 0xC3A, 0xC3F)
    top = np.bincount(keys).max() / len(keys)
    assert abs(top - lut[0]) < 0.003


def test_int64_generators(oracle):
    b = oracle.gen_build_i64(5000, 0x5EED)
    assert len(np.unique(b)) == 5000
    p = oracle.gen_probe_i64(20000, 5000, 0x5EED, 0x77, 0.3)
    assert abs(np.isin(p, b).mean() - 0.3) < 0.02


# ----------------------------------------------------------------- phases
def test_histogram_digit_formulas(oracle):
    rng = np.random.default_rng(7)
    keys = rng.integers(-2**31, 2**31 - 1, 20000).astype(np.int32)
    b1, b2 = 6, 5
    # reference pass-1 digit: ((P1 - 1) << b2 & key) >> b2 evaluated on int32
    ref_d1 = ((np.int64((1 << b1) - 1) << b2) & (keys.astype(np.int64) & 0xFFFFFFFF)) >> b2
    assert np.array_equal(oracle.histogram(keys, b2, b1), np.bincount(ref_d1, minlength=1 << b1))
    ref_d2 = keys.astype(np.int64) & ((1 << b2) - 1)
    assert np.array_equal(oracle.histogram(keys, 0, b2), np.bincount(ref_d2, minlength=1 << b2))
    h = np.zeros(1 << b1, dtype=np.int64)
    oracle.lib().pjo_histogram(keys.ctypes.data_as(oracle.c_vp), len(keys), 4, b2, b1, h.ctypes.data_as(oracle.c_vp))
    assert np.array_equal(h, oracle.histogram(keys, b2, b1))
    fine = oracle.histogram(keys, 0, b1 + b2)
    assert np.array_equal(oracle.fold_histogram(fine, 1 << b2), oracle.histogram(keys, b2, b1))
    assert fine.sum() == len(keys)


def test_scans_closed_form(oracle):
    x = np.array([3, 0, 5, 1, 0, 2])
    assert oracle.exclusive_scan(x).tolist() == [0, 3, 3, 8, 9, 9]
    assert oracle.inclusive_scan(x).tolist() == [3, 3, 8, 9, 9, 11]


def emulate_reference_block_assignment(offsets, total, base):
    """Loop-by-loop emulation of get_number_of_blocks_each_partition +
    InclusiveSum + populate_index_array + InclusiveSum
    (src/skewed_data_handler.cuh:20-91)."""
    P = len(offsets)
    blocks = []
    for p in range(P):
        end = total if p == P - 1 else offsets[p + 1]
        size = end - offsets[p]
        blocks.append(1 if size == 0 else (size - 1) // base + 1)
    scanned = []
    run = 0
    for v in blocks + [0]:
        run += v
        scanned.append(run)
    nblk = scanned[P]
    alt = [0] * (nblk + 1)
    for p in range(P):
        if scanned[p] < len(alt):
            alt[scanned[p]] = 1
    pidx = []
    run = 0
    for b in range(nblk):
        run += alt[b]
        pidx.append(run)
    return scanned[:P], pidx, nblk


def test_block_assignment_closed_form(oracle):
    rng = np.random.default_rng(8)
    for _ in range(20):
        sizes = rng.integers(0, 40, rng.integers(1, 30))
        offsets = oracle.exclusive_scan(sizes)
        total = int(sizes.sum())
        base = int(rng.integers(1, 12))
        scanned, pidx, nblk = oracle.assign_blocks(offsets, total, base)
        e_scanned, e_pidx, e_nblk = emulate_reference_block_assignment(offsets.tolist(), total, base)
        assert scanned.tolist() == e_scanned and nblk == e_nblk
        assert pidx.tolist() == e_pidx
        # chunks tile every partition exactly
        covered = np.zeros(total, dtype=int)
        for blk in range(nblk):
            s, e = oracle.block_chunk(offsets, total, base, scanned, pidx, blk)
            assert e - s <= base
            covered[s:e] += 1
        assert np.all(covered == 1)


def test_partition_pass_restatement(oracle):
    rng = np.random.default_rng(10)
    keys = rng.integers(-2**31, 2**31 - 1, 5000).astype(np.int32)
    pay = rng.integers(0, 10**6, 5000).astype(np.int32)
    for shift, bits in [(0, 4), (3, 7), (8, 8)]:
        ko = np.empty_like(keys)
        po = np.empty_like(pay)
        off = np.empty((1 << bits) + 1, dtype=np.int64)
        oracle.lib().pjo_partition_pass(keys.ctypes.data_as(oracle.c_vp), pay.ctypes.data_as(oracle.c_vp), len(keys),
                                        4, shift, bits, ko.ctypes.data_as(oracle.c_vp), po.ctypes.data_as(oracle.c_vp),
                                        off.ctypes.data_as(oracle.c_vp))
        oracle.check_partitioned(keys, pay, ko, po, off, shift, bits)


def test_partition_checker_catches_planted_mistake(oracle):
    rng = np.random.default_rng(11)
    keys = rng.integers(0, 1000, 400).astype(np.int32)
    pay = np.arange(400, dtype=np.int32)
    ko = np.empty_like(keys)
    po = np.empty_like(pay)
    off = np.empty(17, dtype=np.int64)
    oracle.lib().pjo_partition_pass(keys.ctypes.data_as(oracle.c_vp), pay.ctypes.data_as(oracle.c_vp), 400, 4, 0, 4,
                                    ko.ctypes.data_as(oracle.c_vp), po.ctypes.data_as(oracle.c_vp),
                                    off.ctypes.data_as(oracle.c_vp))
    bad = po.copy()
    bad[0], bad[-1] = bad[-1], bad[0]  # planted: swap payloads across partitions
    with pytest.raises(AssertionError):
        oracle.check_partitioned(keys, pay, ko, bad, off, 0, 4)
    badk = ko.copy()
    badk[0] ^= 1  # planted: key moved to a foreign digit
    with pytest.raises(AssertionError):
        oracle.check_partitioned(keys, pay, badk, po, off, 0, 4)
