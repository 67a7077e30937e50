"""Learned planner: host-side plan decoding and fallback rules (CPU), and the
sketch, exact candidate counts and MLP kernels against numpy / PyTorch fp32
references (GPU).  The planner only picks a plan: joins under the learned plan
and under the plain plan must give the same pairs."""
import ctypes

import numpy as np
import pytest

H = 256
F = 1030
O = 81
HOFF = 17  # 2 * 7 bits classes + 3 pass classes


def _logits(b1=2, b2=2, passes=1, heavy=()):
    lg = np.full(O, -5.0, dtype=np.float32)
    lg[b1] = 5.0
    lg[7 + b2] = 5.0
    lg[14 + passes] = 5.0
    for j in heavy:
        lg[HOFF + j] = 5.0
    return lg


def _cands(n_valid=10, heavy=(), n_probe=1 << 22, border=()):
    """Candidates with exact counts clearly heavy (2T, T = |S| / 1024), clearly
    light (below T / 2) or, for `border`, T (where either prediction passes)."""
    T = n_probe // 1024
    cc = np.zeros(64, dtype=np.uint32)
    cc[:n_valid] = 7
    ex = np.zeros(64, dtype=np.uint64)
    for j in range(n_valid):
        ex[j] = 2 * T if j in heavy else (T if j in border else T // 2 - 1)
    return cc, ex


def test_shapes_and_capacity(pj):
    from partitioned_join_amd import planner
    assert planner.param_count() == 1030 * 256 + 256 + 256 * 256 + 256 + 256 * 81 + 81
    assert planner.chunk_capacity(4) == 2783 and planner.chunk_capacity(8) == 2047
    assert planner.chunk_capacity(3) == 0


def test_decode_takes_learned_plan_when_heavy_sets_agree(pj):
    from partitioned_join_amd import planner
    n_b, n_p = 1 << 26, 1 << 28
    cc, ex = _cands(12, heavy=(0, 1, 2), n_probe=n_p)
    plan, fb = planner.decode(_logits(2, 2, 1, heavy=(0, 1, 2)), cc, ex, n_b, n_p, 4)
    assert not fb and plan.source == 1
    assert (plan.passes, plan.bits1, plan.bits2) == (2, 8, 8)


def test_decode_falls_back_on_any_heavy_disagreement(pj):
    from partitioned_join_amd import planner
    n_b, n_p = 1 << 26, 1 << 28
    plain = pj.plan_default(n_b, n_p)
    cc, ex = _cands(12, heavy=(0, 1, 2), n_probe=n_p)
    for heavy_pred in [(0, 1), (0, 1, 2, 3), ()]:
        plan, fb = planner.decode(_logits(2, 2, 1, heavy=heavy_pred), cc, ex, n_b, n_p, 4)
        assert fb and plan.source == 0
        assert (plan.passes, plan.bits1, plan.bits2) == (plain.passes, plain.bits1, plain.bits2)
    # a candidate with no sample occurrences is not part of the comparison
    cc2 = cc.copy()
    cc2[3] = 0
    plan, fb = planner.decode(_logits(2, 2, 1, heavy=(0, 1, 2, 3)), cc2, ex, n_b, n_p, 4)
    assert not fb
    # the band edges: T/2 - 1 rows predicted heavy fails, T/2 passes; 2T - 1 missed passes, 2T fails
    T = n_p // 1024
    for count, pred, ok in [(T // 2 - 1, True, False), (T // 2, True, True), (2 * T - 1, False, True),
                            (2 * T, False, False)]:
        ex2 = ex.copy()
        ex2[5] = count
        heavy_pred = (0, 1, 2, 5) if pred else (0, 1, 2)
        plan, fb = planner.decode(_logits(2, 2, 1, heavy=heavy_pred), cc, ex2, n_b, n_p, 4)
        assert fb == (not ok), (count, pred)


def test_decode_accepts_either_answer_near_the_threshold(pj):
    from partitioned_join_amd import planner
    n_b, n_p = 1 << 26, 1 << 28
    cc, ex = _cands(12, heavy=(0,), border=(1, 2), n_probe=n_p)
    for heavy_pred in [(0,), (0, 1), (0, 1, 2)]:
        plan, fb = planner.decode(_logits(2, 2, 1, heavy=heavy_pred), cc, ex, n_b, n_p, 4)
        assert not fb


def test_heavy_targets_follow_the_band(pj):
    from partitioned_join_amd import planner
    n = 1 << 20
    T = n // 1024
    t = planner.heavy_targets(np.array([3, 3, 3, 3, 0]), np.array([2 * T, 2 * T - 1, T // 2, T // 2 - 1, 5 * T]), n)
    assert t.tolist() == [1.0, -1.0, -1.0, 0.0, -1.0]


def test_decode_rejects_plans_whose_chunks_do_not_fit(pj):
    from partitioned_join_amd import planner
    n_b, n_p = 1 << 26, 1 << 27
    cc, ex = _cands(0, n_probe=n_p)
    plan, fb = planner.decode(_logits(2, 0, 0), cc, ex, n_b, n_p, 4)  # one pass of 8 bits for 64M rows
    assert fb and plan.source == 0
    plan, fb = planner.decode(_logits(2, 0, 0), cc, ex, 100000, n_p, 4)  # 586 rows per partition: fits
    assert not fb and (plan.passes, plan.bits1, plan.bits2) == (1, 8, 0)


def test_decode_clamps_three_passes_and_total_bits(pj):
    from partitioned_join_amd import planner
    cc, ex = _cands(0)
    plan, fb = planner.decode(_logits(1, 1, 2), cc, ex, 1 << 20, 1 << 22, 4)  # 3 passes of 7 + 7
    assert not fb and (plan.passes, plan.bits1, plan.bits2) == (2, 7, 7)
    plan, fb = planner.decode(_logits(6, 6, 1), cc, ex, 1 << 20, 1 << 22, 4)  # 12 + 12 bits
    assert not fb and (plan.passes, plan.bits1, plan.bits2) == (2, 12, 4)


def test_candidate_plans_and_join_specs(pj):
    from partitioned_join_amd import planner
    big = planner.candidate_plans(1 << 26, 4)
    assert big == [(2, 8, 8)]
    small = planner.candidate_plans(1 << 20, 4)
    assert (1, 10, 0) in small and (2, 6, 6) in small and len(small) == 12
    assert (1, 9, 0) not in small  # 1.5 * 2^20 / 2^9 = 3072 rows per chunk > 2783
    assert planner.candidate_plans(1 << 26, 8) == [(2, 8, 8)]
    specs = [planner.join_spec(i) for i in range(200)]
    assert specs[5] == planner.join_spec(5)
    for s in specs:
        assert (1 << 20) <= s["n_build"] <= (1 << 26)
        assert s["n_build"] <= s["n_probe"] <= 16 * s["n_build"] + 1
        assert s["key_bytes"] in (4, 8) and 0 <= s["theta"] <= 1.5 and 0.1 <= s["match"] <= 1.0
    assert {s["key_bytes"] for s in specs} == {4, 8}


# ------------------------------------------------------------------ GPU
@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def sketch_numpy(oracle, probe, n_build):
    """Restatement of the sketch: positions ((2j+1)n) // 8192, hash bins
    murmur3 & 1023 sorted by count, runs of equal sampled keys ordered by
    count (desc) then key (as unsigned)."""
    n = probe.size
    kb = probe.itemsize
    pos = ((2 * np.arange(4096, dtype=np.int64) + 1) * n) // 8192
    smp = probe[pos]
    u = smp.astype(np.int64) & (0xFFFFFFFF if kb == 4 else -1)
    bins = np.zeros(1024, dtype=np.int64)
    for k in smp.tolist():
        bins[oracle.murmur3_key(int(k), kb) & 1023] += 1
    feat = np.zeros(1030, dtype=np.float32)
    feat[:1024] = np.sort(bins)[::-1] / 4096.0
    uk = u.view(np.uint64) if kb == 8 else u.astype(np.uint64)
    vals, cnts = np.unique(uk, return_counts=True)
    order = np.lexsort((vals, -cnts))
    top_keys = np.zeros(64, dtype=np.uint64)
    top_cnt = np.zeros(64, dtype=np.int64)
    m = min(64, len(order))
    top_keys[:m] = vals[order[:m]]
    top_cnt[:m] = cnts[order[:m]]
    nb, ns = n_build + 1.0, n + 1.0
    feat[1024:] = [np.log2(nb) / 32, np.log2(ns) / 32, np.log2(ns / nb) / 8, kb / 8, cnts.max() / 4096,
                   float((cnts.astype(np.float64) ** 2).sum() / 4096 ** 2)]
    return feat, top_keys, top_cnt


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_sketch_and_candidate_counts(pj, oracle, torch, dtype):
    from partitioned_join_amd import planner
    rng = np.random.default_rng(31)
    n = 1_000_003
    hot = rng.integers(-2**31, 2**31 - 1, 40).astype(dtype)
    probe = rng.integers(-2**31, 2**31 - 1, n).astype(dtype)
    idx = rng.random(n) < 0.3
    probe[idx] = hot[rng.integers(0, 40, idx.sum()) ** 2 % 40]
    build = np.arange(5000, dtype=dtype)
    d = lambda a: torch.from_numpy(a).cuda()
    feat, cand, cnt = planner.sketch(d(build), d(probe))
    ef, ek, ec = sketch_numpy(oracle, probe, build.size)
    f = feat.cpu().numpy()
    assert np.array_equal(f[:1024], ef[:1024])
    assert np.allclose(f[1024:], ef[1024:], rtol=1e-6, atol=1e-7)
    got_k = cand.cpu().numpy().view(np.uint64)
    if dtype == np.int32:
        got_k = got_k & np.uint64(0xFFFFFFFF)
    assert np.array_equal(cnt.cpu().numpy().astype(np.int64), ec)
    assert np.array_equal(got_k[ec > 0], ek[ec > 0])
    exact = planner.candidate_counts(d(probe), cand, cnt).cpu().numpy()
    pu = probe.astype(np.int64).view(np.uint64) if dtype == np.int64 else (probe.astype(np.int64) & 0xFFFFFFFF).astype(np.uint64)
    for j in range(64):
        want = int((pu == ek[j]).sum()) if ec[j] > 0 else 0
        assert exact[j] == want, j
    # a column that starts one key in is not 16-byte aligned: the per-key kernel
    shifted = planner.candidate_counts(d(probe)[1:], cand, cnt).cpu().numpy()
    for j in range(64):
        want = int((pu[1:] == ek[j]).sum()) if ec[j] > 0 else 0
        assert shifted[j] == want, j
    # empty probe: no candidates, zero features
    feat0, _, cnt0 = planner.sketch(d(build), d(probe[:0]))
    assert float(feat0[:1024].abs().sum()) == 0.0 and int(cnt0.sum()) == 0


def torch_forward(torch, w, x):
    W1 = w[:F * H].view(F, H); b1 = w[F * H:F * H + H]
    o = F * H + H
    W2 = w[o:o + H * H].view(H, H); b2 = w[o + H * H:o + H * H + H]
    o += H * H + H
    W3 = w[o:o + H * O].view(H, O); b3 = w[o + H * O:o + H * O + O]
    h1 = torch.relu(x @ W1 + b1)
    h2 = torch.relu(h1 @ W2 + b2)
    return h2 @ W3 + b3


def torch_loss(torch, out, labels, heavy):
    import torch.nn.functional as Fn
    B = out.shape[0]
    tot = 0.0
    for b in range(B):
        l = Fn.cross_entropy(out[b:b + 1, 0:7], labels[b:b + 1, 0].long(), reduction="sum")
        if labels[b, 1] >= 0:
            l = l + Fn.cross_entropy(out[b:b + 1, 7:14], labels[b:b + 1, 1].long(), reduction="sum")
        l = l + Fn.cross_entropy(out[b:b + 1, 14:17], labels[b:b + 1, 2].long(), reduction="sum")
        m = heavy[b] >= 0
        if m.any():
            l = l + Fn.binary_cross_entropy_with_logits(out[b, HOFF:][m], heavy[b][m], reduction="sum") / 64.0
        tot = tot + l
    return tot / B


@pytest.mark.gpu
def test_forward_matches_torch_fp32(pj, torch):
    from partitioned_join_amd import planner
    w = planner.init_weights(3)
    x = torch.rand((7, F), device="cuda")
    got = planner.forward(w, x)
    ref = torch_forward(torch, w, x)
    assert torch.allclose(got, ref, rtol=1e-4, atol=1e-4), (got - ref).abs().max()


@pytest.mark.gpu
def test_train_step_matches_torch_adam(pj, torch):
    from partitioned_join_amd import planner
    g = torch.Generator().manual_seed(5)
    B = 9
    w0 = planner.init_weights(4)
    x = torch.rand((B, F), generator=g).cuda()
    labels = torch.stack([torch.randint(0, 7, (B,), generator=g), torch.randint(-1, 7, (B,), generator=g),
                          torch.randint(0, 3, (B,), generator=g)], 1).int().cuda()
    heavy = torch.randint(-1, 2, (B, 64), generator=g).float().cuda()
    # reference: torch autograd + Adam on the same loss
    wr = w0.clone().requires_grad_(True)
    opt = torch.optim.Adam([wr], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    lref = torch_loss(torch, torch_forward(torch, wr, x), labels, heavy)
    opt.zero_grad()
    lref.backward()
    opt.step()
    w = w0.clone()
    m = torch.zeros_like(w)
    v = torch.zeros_like(w)
    loss = planner.train_step(w, m, v, 1, 1e-3, x, labels, heavy)
    assert abs(float(loss.item()) - float(lref.item())) < 1e-4 * max(1.0, abs(float(lref.item())))
    # Adam's first step moves every weight with a nonzero gradient by ~lr
    assert torch.allclose(w, wr.detach(), atol=2e-5), (w - wr.detach()).abs().max()
    # a few more steps on the same batch lower the loss
    for s in range(2, 40):
        last = planner.train_step(w, m, v, s, 1e-3, x, labels, heavy)
    assert float(last.item()) < 0.5 * float(loss.item())


@pytest.mark.gpu
def test_plan_learned_agreement_fallback_and_same_pairs(pj, oracle, torch):
    from partitioned_join_amd import planner
    G = pj.generators
    bk = G.build_unique_i32(1 << 20, 1 << 20, 0x71)
    pz = G.probe_zipf_i32(1 << 22, 1 << 20, 1.2, 0x72, 0x73)
    # all-zero weights: logits 0, nothing predicted heavy, but Zipf(1.2) has heavy keys
    w = torch.zeros(planner.param_count(), device="cuda")
    plan, fb = planner.plan_learned(bk, pz, w)
    plain = pj.plan_default(bk.numel(), pz.numel())
    assert fb and (plan.passes, plan.bits1, plan.bits2) == (plain.passes, plain.bits1, plain.bits2)
    # weights whose output is b3 alone, set from the exact heavy set: the learned plan is taken
    _, cand, cnt = planner.sketch(bk, pz)
    exact = planner.candidate_counts(pz, cand, cnt).cpu().numpy()
    heavy = (exact * 1024 >= pz.numel()) & (cnt.cpu().numpy() > 0)
    assert heavy.any()
    b3 = np.full(O, -4.0, dtype=np.float32)
    b3[1] = 4.0      # pass 1: 7 bits
    b3[7 + 1] = 4.0  # pass 2: 7 bits
    b3[14 + 1] = 4.0  # two passes
    b3[HOFF:] = np.where(heavy, 4.0, -4.0)
    w[-O:] = torch.from_numpy(b3).cuda()
    plan2, fb2 = planner.plan_learned(bk, pz, w)
    assert not fb2 and plan2.source == 1 and (plan2.passes, plan2.bits1, plan2.bits2) == (2, 7, 7)
    a = oracle.sort_pairs(pj.partitioned_join(bk, pz, plan=plan2).cpu().numpy())
    b = oracle.sort_pairs(pj.partitioned_join(bk, pz, plan=plain).cpu().numpy())
    assert len(a) == pz.numel() and np.array_equal(a, b)
    # weights predicting the plain plan itself (one pass of 9 bits) and no heavy key: there is
    # nothing for the heavy-set check to choose between, so the plan is taken without it
    assert (plain.passes, plain.bits1) == (1, 9)
    b3 = np.full(O, -4.0, dtype=np.float32)
    b3[9 - 6] = 4.0   # pass 1: 9 bits
    b3[14 + 0] = 4.0  # one pass
    w[-O:] = torch.from_numpy(b3).cuda()
    plan3, fb3 = planner.plan_learned(bk, pz, w)
    assert not fb3 and plan3.source == 1 and (plan3.passes, plan3.bits1, plan3.bits2) == (1, 9, 0)


@pytest.mark.gpu
@pytest.mark.parametrize("i", [2, 3, 4, 5, 6, 10, 11])
def test_trained_weights_plan_joins_equal_to_oracle(pj, oracle, torch, i):
    """The committed trained weights plan training-distribution joins (key
    width, Zipf theta and match rate of planner.join_spec(i), sizes scaled to at
    most 1M x 4M for the oracle); whatever plan the model and its fallback pick,
    the pairs equal the oracle's join."""
    from partitioned_join_amd import planner
    w = planner.load_weights()
    assert w is not None, "trained weights are committed with the package"
    spec = planner.join_spec(i)
    scale = max(1.0, spec["n_build"] / (1 << 20), spec["n_probe"] / (1 << 22))
    spec = dict(spec, n_build=max(1, int(spec["n_build"] / scale)), n_probe=max(1, int(spec["n_probe"] / scale)))
    bk, pk = planner.make_join(spec)
    plan, fallback = planner.plan_learned(bk, pk, w)
    print(f"join {i}: {spec['n_build']} x {spec['n_probe']}, {spec['key_bytes']}-byte keys, theta "
          f"{spec['theta']:.2f}: plan {plan.as_dict()}, fallback {fallback}", flush=True)
    assert plan.source == (0 if fallback else 1)
    got = pj.partitioned_join(bk, pk, plan=plan, out_capacity=pk.numel() + 1)
    exp = oracle.inner_join_cpu(bk.cpu().numpy(), pk.cpu().numpy())
    assert np.array_equal(oracle.sort_pairs(got.cpu().numpy()), oracle.sort_pairs(exp))
