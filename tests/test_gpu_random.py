"""Seeded random joins on the GPU against the oracle: key width, key range
(so duplicates and misses vary), table sizes (ragged, not powers of two),
payload columns, plan (passes, radix bits, chunk factor, probe slice) and the
unique-key promise are all drawn per case.  Each case is small enough for
the oracle's CPU join; the sorted (build_idx, probe_idx) multisets must be
equal."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def _case(seed):
    r = np.random.default_rng([0x5EED, seed])
    kb = 4 if r.random() < 0.5 else 8
    nb = int(r.integers(1, 150_000))
    ns = int(r.integers(1, 400_000))
    span = int(r.choice([nb // 3 + 1, nb, 4 * nb + 7, 2**31 - 1]))
    dt = np.int32 if kb == 4 else np.int64
    lo = -span // 2
    bk = (r.integers(0, span, nb) + lo).astype(dt)
    pk = (r.integers(0, span, ns) + lo).astype(dt)
    if kb == 8:  # exercise the high key word too
        bk = (bk.astype(np.int64) * (2**33 + 1)).astype(np.int64)
        pk = (pk.astype(np.int64) * (2**33 + 1)).astype(np.int64)
    bp = r.integers(-2**31, 2**31 - 1, nb).astype(np.int32) if r.random() < 0.4 else None
    pp = r.integers(-2**31, 2**31 - 1, ns).astype(np.int32) if r.random() < 0.4 else None
    passes = int(r.integers(1, 3))
    b1 = int(r.integers(1, 13))
    b2 = int(r.integers(1, 13)) if passes == 2 else 0
    if b1 + b2 > 16:
        b2 = 16 - b1 if passes == 2 else 0
        if passes == 2 and b2 < 1:
            b1, b2 = 8, 8
    f = float(r.choice([1.0, 1.5, 3.0]))
    # the largest build chunk must fit the LDS table: raise the bits if needed
    cap = 2047 if kb == 8 else 2783
    while f * nb / 2 ** (b1 + b2) > cap and b1 + b2 < 16:
        if passes == 1 and b1 < 12:
            b1 += 1
        elif passes == 2 and b2 < 12 and b1 + b2 < 16:
            b2 += 1
        else:
            break
    plan = {"passes": passes, "bits1": b1, "bits2": b2, "max_partition_size_factor": f,
            "probe_slice": int(r.choice([0, 0, 1000, 5000]))}
    uniq = len(np.unique(bk)) == nb
    plan["unique_build_keys"] = int(uniq and r.random() < 0.5)
    return kb, bk, pk, bp, pp, plan


@pytest.mark.parametrize("seed", range(24))
def test_random_join_matches_oracle(pj, oracle, torch, seed):
    kb, bk, pk, bp, pp, plan = _case(seed)
    d = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda()
    print(f"seed {seed}: key bytes {kb}, |R| {len(bk)}, |S| {len(pk)}, plan {plan}", flush=True)
    got = pj.partitioned_join(d(bk), d(pk), d(bp), d(pp), plan=plan)
    exp = oracle.inner_join_cpu(bk, pk, bp, pp, uniq_keys=False)
    g = oracle.sort_pairs(got.cpu().numpy())
    e = oracle.sort_pairs(exp)
    assert len(g) == len(e), (seed, len(g), len(e))
    assert np.array_equal(g, e), seed
