"""The allocation-free entry point (pj_partitioned_join_noalloc, static device
scratch) against pj_partitioned_join_async with a caller workspace: two calls
in a row on the same GPU, the second with smaller tables, equal pair counts and
equal sorted pairs (and both equal to the oracle's join on the second)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def _async_join(pj, torch, bk, pk, bp=None, pp=None):
    out = torch.empty((max(pk.numel(), 1), 2), dtype=torch.int32, device="cuda")
    ws = torch.empty(pj.workspace_size(bk.numel(), pk.numel(), bk.element_size()), dtype=torch.uint8, device="cuda")
    d = pj.partitioned_join_async(bk, pk, out, bp, pp, workspace=ws)
    torch.cuda.synchronize()
    n = int(d.item())
    return out[:n].cpu().numpy(), n


def _noalloc_join(pj, torch, bk, pk, bp=None, pp=None):
    out = torch.full((max(pk.numel(), 1), 2), -5, dtype=torch.int32, device="cuda")
    d = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    pj.partitioned_join_noalloc(bk, pk, out, d, bp, pp)
    torch.cuda.synchronize()
    n = int(d.item())
    return out[:n].cpu().numpy(), n


def test_noalloc_equals_async_two_calls_second_smaller(pj, oracle, torch):
    G = pj.generators
    # call 1: 1M unique int32 build keys x 4M probes (every probe matches once)
    bk1, pk1 = G.workload_config1()
    print("call 1: 1M x 4M int32", flush=True)
    a1, n1 = _async_join(pj, torch, bk1, pk1)
    b1, m1 = _noalloc_join(pj, torch, bk1, pk1)
    assert n1 == m1 == 1 << 22
    assert np.array_equal(oracle.sort_pairs(a1), oracle.sort_pairs(b1))
    # call 2: smaller, int64 keys with duplicates and payload columns
    rng = np.random.default_rng(0x7A)
    bk = rng.integers(0, 30000, 50000).astype(np.int64) * (2**33 + 1)
    pk = rng.integers(0, 30000, 120000).astype(np.int64) * (2**33 + 1)
    bp = rng.integers(-2**31, 2**31 - 1, bk.size).astype(np.int32)
    pp = rng.integers(-2**31, 2**31 - 1, pk.size).astype(np.int32)
    d = lambda a: torch.from_numpy(a).cuda()
    print("call 2: 50K x 120K int64 with duplicates and payloads", flush=True)
    exp = oracle.sort_pairs(oracle.inner_join_cpu(bk, pk, bp, pp, uniq_keys=False))
    cap_out = torch.empty((len(exp) + 16, 2), dtype=torch.int32, device="cuda")
    dc = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    pj.partitioned_join_noalloc(d(bk), d(pk), cap_out, dc, d(bp), d(pp))
    torch.cuda.synchronize()
    m2 = int(dc.item())
    assert m2 == len(exp)
    assert np.array_equal(oracle.sort_pairs(cap_out[:m2].cpu().numpy()), exp)
    got_async = pj.partitioned_join(d(bk), d(pk), d(bp), d(pp), out_capacity=len(exp) + 16)
    assert np.array_equal(oracle.sort_pairs(got_async.cpu().numpy()), exp)


def test_noalloc_rejects_joins_beyond_its_scratch(pj, torch):
    big_b = torch.empty(1 << 24, dtype=torch.int32, device="cuda")
    big_p = torch.empty(1 << 26, dtype=torch.int32, device="cuda")
    out = torch.empty((16, 2), dtype=torch.int32, device="cuda")
    d = torch.zeros(1, dtype=torch.int64, device="cuda")
    with pytest.raises(pj.PjError) as e:
        pj.partitioned_join_noalloc(big_b, big_p, out, d)
    assert e.value.status == 3 and "allocation-free" in str(e.value)
