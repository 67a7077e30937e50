"""Learned join planner: a small FP32 MLP over a sampled key sketch that picks
radix bits and pass count, with the plain plan as the fallback.

Inference and training run in libpj.so's own HIP kernels (pj_planner_*); this
module holds the weights, the training-set generator and the training loop.

The model reads a 1030-float sketch of the probe column:
- a 4096-key sample's 1024-bin hash histogram, with bins sorted by count;
- six size and skew scalars.

It predicts three things:
- the radix bits of each pass (6..12);
- the pass count (1..3; three passes run as two);
- for each of the 64 most frequent sampled keys, whether it holds at least
  1/1024 of the probe rows.

The library takes the learned plan only when that predicted heavy set agrees
with the exact candidate counts over the whole probe column, within a band: a
key predicted heavy must hold at least T/2 rows and a key predicted light
fewer than 2T, for T = |S|/1024. Otherwise it returns the plain
histogram-driven plan. Either way the join result is the same.

Training data follows the planner training/eval workload: seeded joins with
|R| log-uniform in [1M, 64M], |S|/|R| in [1, 16], int32 or int64 keys, Zipf
theta in [0, 1.5] and match rate in [0.1, 1]. Labels come from the exact
candidate counts and from the fastest of the candidate plans, timed with the
library's phase events.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time
from typing import Optional

import numpy as np

from . import _lib
from ._lib import PjPlan, PjTable

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_PATH = os.path.join(HERE, "planner_weights.bin")
META_PATH = os.path.join(HERE, "planner_weights.json")

FEATURES = 1030
CANDIDATES = 64
LOGITS = 81
HIDDEN = 256
BITS_LO, BITS_CLASSES = 6, 7
PASS_CLASSES = 3


def _t():
    import torch
    return torch


def _s(stream=None):
    torch = _t()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _L():
    return _lib.load()


def param_count() -> int:
    return int(_L().pj_planner_param_count())


def chunk_capacity(key_bytes: int) -> int:
    return int(_L().pj_join_chunk_capacity(key_bytes))


def init_weights(seed: int = 0, device="cuda"):
    """He-initialised weights in the library's layout (W1, b1, W2, b2, W3, b3)."""
    torch = _t()
    g = torch.Generator().manual_seed(seed)
    parts = []
    for fan_in, fan_out in [(FEATURES, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, LOGITS)]:
        parts.append(torch.randn(fan_in * fan_out, generator=g) * math.sqrt(2.0 / fan_in))
        parts.append(torch.zeros(fan_out))
    w = torch.cat(parts).float()
    assert w.numel() == param_count()
    return w.to(device)


def load_weights(path: str = WEIGHTS_PATH, device="cuda"):
    """The committed trained weights, or None when none were trained."""
    torch = _t()
    if not os.path.exists(path):
        return None
    w = np.fromfile(path, dtype="<f4")
    if w.size != param_count():
        raise ValueError(f"{path}: {w.size} floats, the model has {param_count()}")
    return torch.from_numpy(w.copy()).to(device)


def save_weights(w, meta: dict, path: str = WEIGHTS_PATH, meta_path: str = META_PATH) -> None:
    w.detach().float().cpu().numpy().astype("<f4").tofile(path)
    with open(meta_path, "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


def _table(keys) -> PjTable:
    return PjTable(keys.data_ptr(), None, keys.numel(), keys.element_size(), 0)


def sketch(build_keys, probe_keys, stream=None):
    """(features[1030] float32, candidate keys[64] int64, sample counts[64] int32)."""
    torch = _t()
    dev = probe_keys.device
    feat = torch.empty(FEATURES, dtype=torch.float32, device=dev)
    cand = torch.empty(CANDIDATES, dtype=torch.int64, device=dev)
    cnt = torch.empty(CANDIDATES, dtype=torch.int32, device=dev)
    _lib.check(_L().pj_planner_sketch(ctypes.byref(_table(build_keys)), ctypes.byref(_table(probe_keys)),
                                      feat.data_ptr(), cand.data_ptr(), cnt.data_ptr(), _s(stream)))
    return feat, cand, cnt


def candidate_counts(probe_keys, cand, cand_counts, stream=None):
    """Exact occurrences of each candidate key in the probe column (int64[64])."""
    torch = _t()
    out = torch.empty(CANDIDATES, dtype=torch.int64, device=probe_keys.device)
    _lib.check(_L().pj_planner_candidate_counts(ctypes.byref(_table(probe_keys)), cand.data_ptr(),
                                                cand_counts.data_ptr(), out.data_ptr(), _s(stream)))
    return out


def _scratch(batch: int, device):
    torch = _t()
    nbytes = int(_L().pj_planner_scratch_bytes(batch))
    return torch.empty(nbytes, dtype=torch.uint8, device=device)


def forward(w, x, stream=None):
    """Logits [batch, 81] for features x [batch, 1030]."""
    torch = _t()
    x = x.contiguous().float()
    batch = x.shape[0]
    out = torch.empty((batch, LOGITS), dtype=torch.float32, device=x.device)
    scr = _scratch(batch, x.device)
    _lib.check(_L().pj_planner_forward(w.data_ptr(), x.data_ptr(), batch, out.data_ptr(), scr.data_ptr(),
                                       scr.numel(), _s(stream)))
    return out


def train_step(w, m, v, step: int, lr: float, x, labels, heavy, stream=None):
    """One Adam step on a batch; returns the mean loss (device scalar)."""
    torch = _t()
    batch = x.shape[0]
    loss = torch.empty(1, dtype=torch.float32, device=x.device)
    scr = _scratch(batch, x.device)
    _lib.check(_L().pj_planner_train_step(w.data_ptr(), m.data_ptr(), v.data_ptr(), int(step), float(lr),
                                          x.contiguous().data_ptr(), labels.contiguous().data_ptr(),
                                          heavy.contiguous().data_ptr(), batch, loss.data_ptr(), scr.data_ptr(),
                                          scr.numel(), _s(stream)))
    return loss


def decode(logits, cand_counts, exact_counts, n_build: int, n_probe: int, key_bytes: int):
    """Host decode of one sample (pj_plan_decode): (plan, fallback)."""
    lg = np.ascontiguousarray(np.asarray(logits, dtype=np.float32).reshape(-1))
    cc = np.ascontiguousarray(np.asarray(cand_counts, dtype=np.uint32).reshape(-1))
    ex = np.ascontiguousarray(np.asarray(exact_counts, dtype=np.uint64).reshape(-1))
    plan = PjPlan()
    fb = ctypes.c_int32(0)
    bt = PjTable(None, None, int(n_build), int(key_bytes), 0)
    pt = PjTable(None, None, int(n_probe), int(key_bytes), 0)
    _lib.check(_L().pj_plan_decode(lg.ctypes.data, cc.ctypes.data, ex.ctypes.data, ctypes.byref(bt),
                                   ctypes.byref(pt), ctypes.byref(plan), ctypes.byref(fb)))
    return plan, bool(fb.value)


def plan_learned(build_keys, probe_keys, w=None, stream=None):
    """The learned plan for this join (pj_plan_learned), or the plain plan
    when the prediction disagrees with the exact heavy keys: (plan, fallback)."""
    torch = _t()
    if w is None:
        w = load_weights(device=probe_keys.device)
        if w is None:
            raise FileNotFoundError(f"no trained planner weights at {WEIGHTS_PATH}")
    ws = torch.empty(int(_L().pj_plan_learned_workspace_size()), dtype=torch.uint8, device=probe_keys.device)
    plan = PjPlan()
    fb = ctypes.c_int32(0)
    _lib.check(_L().pj_plan_learned(ctypes.byref(_table(build_keys)), ctypes.byref(_table(probe_keys)),
                                    w.data_ptr(), ctypes.byref(plan), ctypes.byref(fb), ws.data_ptr(), ws.numel(),
                                    _s(stream)))
    return plan, bool(fb.value)


# --------------------------------------------------------------------------- training data
def candidate_plans(n_build: int, key_bytes: int, factor: float = 1.5):
    """Plans the labels choose from: one pass of 8..12 bits, two passes of
    6..8 + 6..8 bits; only plans whose average build chunk fits the LDS table."""
    cap = chunk_capacity(key_bytes)
    plans = [(1, b, 0) for b in range(8, 13)] + [(2, b1, b2) for b1 in (6, 7, 8) for b2 in (6, 7, 8)]
    ok = [p for p in plans if factor * n_build / float(1 << (p[1] + p[2])) <= cap]
    if not ok:
        ok = [max(plans, key=lambda p: p[1] + p[2])]
    return ok


def join_spec(i: int, seed: int = 0x9A):
    """Parameters of training join i (deterministic)."""
    r = np.random.default_rng([seed, i])
    nb = int(round(math.exp(r.uniform(math.log(1 << 20), math.log(1 << 26)))))
    ratio = r.uniform(1.0, 16.0)
    ns = min(int(round(nb * ratio)), (1 << 31) - 1)
    kb = 4 if r.random() < 0.5 else 8
    theta = float(r.uniform(0.0, 1.5))
    match = float(r.uniform(0.1, 1.0))
    return {"i": i, "n_build": nb, "n_probe": ns, "key_bytes": kb, "theta": theta, "match": match,
            "seeds": [int(x) for x in r.integers(1, 2**62, 4)]}


def make_join(spec, device="cuda"):
    """Device build/probe keys of a training join: unique build keys; probe
    keys drawn Zipf(theta) over the build keys (rank -> key by a seeded
    permutation), a (1 - match) fraction replaced by keys absent from R."""
    torch = _t()
    L = _L()
    nb, ns, kb, theta = spec["n_build"], spec["n_probe"], spec["key_bytes"], spec["theta"]
    s0, s1, s2, s3 = spec["seeds"]
    s = _s()
    # Zipf CDF over the build ranks, computed on the device in float64
    w = torch.arange(1, nb + 1, dtype=torch.float64, device=device).pow_(-theta)
    lut = torch.cumsum(w, 0)
    lut /= lut[-1].clone()
    del w
    if kb == 4:
        bk = torch.empty(nb, dtype=torch.int32, device=device)
        _lib.check(L.pj_gen_build_unique_i32(bk.data_ptr(), nb, nb, s0, s))
        pk = torch.empty(ns, dtype=torch.int32, device=device)
        _lib.check(L.pj_gen_probe_zipf_i32(pk.data_ptr(), ns, lut.data_ptr(), nb, s1, s2, s))
    else:
        bk = torch.empty(nb, dtype=torch.int64, device=device)
        _lib.check(L.pj_gen_build_i64(bk.data_ptr(), nb, s0, s))
        pk = torch.empty(ns, dtype=torch.int64, device=device)
        _lib.check(L.pj_gen_probe_zipf_i64(pk.data_ptr(), ns, lut.data_ptr(), nb, s0, s1, s2, s))
    del lut
    g = torch.Generator(device=device).manual_seed(s3 & 0x7FFFFFFF)
    miss = torch.rand(ns, generator=g, device=device) >= spec["match"]
    k = int(miss.sum().item())
    if k:
        if kb == 4:  # build keys are a permutation of [0, nb): keys >= nb never match
            fill = torch.randint(nb, 2 * nb, (k,), generator=g, device=device, dtype=torch.int64).to(torch.int32)
        else:        # 64-bit random keys: a collision with R has probability < 2^-30
            fill = torch.randint(-2**62, 2**62, (k,), generator=g, device=device, dtype=torch.int64)
        pk[miss] = fill
    return bk, pk


def time_plan(bk, pk, plan_tuple, reps: int = 2):
    """Fastest of `reps` timed joins (library phase events, total_ms)."""
    from . import partitioned_join_async, profile_enable, profile_read, workspace_size
    torch = _t()
    passes, b1, b2 = plan_tuple
    plan = {"passes": passes, "bits1": b1, "bits2": b2, "max_partition_size_factor": 1.5, "source": 2}
    ws = torch.empty(workspace_size(bk.numel(), pk.numel(), bk.element_size(), plan), dtype=torch.uint8,
                     device=bk.device)
    cap = pk.numel() + 1024
    out = torch.empty((cap, 2), dtype=torch.int32, device=bk.device)
    partitioned_join_async(bk, pk, out, plan=plan, workspace=ws)  # warm-up
    best = float("inf")
    profile_enable(True)
    for _ in range(reps):
        d = partitioned_join_async(bk, pk, out, plan=plan, workspace=ws)
        best = min(best, profile_read()["total_ms"])
    profile_enable(False)
    n = int(d.item())
    del out, ws
    return best, n


def heavy_targets(cand_counts, exact_counts, n_probe: int):
    """Training targets of the heavy flags, matching the library's agreement
    rule (T = |S| / 1024): 1 for at least 2T rows, 0 below T / 2, -1 (no
    loss) in between, where either answer is accepted, and for missing
    candidates."""
    c = np.asarray(exact_counts, dtype=np.float64)
    t = np.where(c * 512 >= n_probe, 1.0, np.where(c * 2048 < n_probe, 0.0, -1.0))
    return np.where(np.asarray(cand_counts) > 0, t, -1.0).astype(np.float32)


def label_join(spec, device="cuda", reps: int = 2):
    """Features, candidate info, heavy targets and plan timings of one join."""
    torch = _t()
    bk, pk = make_join(spec, device)
    feat, cand, cnt = sketch(bk, pk)
    exact = candidate_counts(pk, cand, cnt)
    cnt_h = cnt.cpu().numpy()
    exact_h = exact.cpu().numpy()
    heavy = heavy_targets(cnt_h, exact_h, pk.numel())
    timings = {}
    counts = set()
    for p in candidate_plans(bk.numel(), bk.element_size()):
        ms, n = time_plan(bk, pk, p, reps)
        timings[p] = ms
        counts.add(n)
    assert len(counts) == 1, f"plans disagree on the pair count: {counts}"
    best = min(timings, key=timings.get)
    passes, b1, b2 = best
    labels = np.array([b1 - BITS_LO, (b2 - BITS_LO) if passes == 2 else -1, passes - 1], dtype=np.int32)
    rec = {"spec": spec, "timings": {f"{p[0]}:{p[1]}:{p[2]}": t for p, t in timings.items()},
           "best": list(best), "pairs": counts.pop()}
    del bk, pk
    torch.cuda.empty_cache()
    return feat.cpu().numpy(), labels, heavy, cnt_h, exact_h, rec


def build_dataset(first: int, count: int, out_prefix: str, device="cuda", log=print):
    """Labels joins first .. first + count - 1 and saves them as .npz + .json."""
    X, Y, H, C, E, recs = [], [], [], [], [], []
    t0 = time.time()
    for i in range(first, first + count):
        f, y, h, c, e, rec = label_join(join_spec(i), device)
        X.append(f); Y.append(y); H.append(h); C.append(c); E.append(e); recs.append(rec)
        if (i - first) % 25 == 0:
            log(f"join {i}: {rec['spec']['n_build']} x {rec['spec']['n_probe']} kb={rec['spec']['key_bytes']} "
                f"theta={rec['spec']['theta']:.2f} best={rec['best']} ({time.time() - t0:.0f}s)")
    np.savez(out_prefix + ".npz", X=np.stack(X), Y=np.stack(Y), H=np.stack(H), C=np.stack(C), E=np.stack(E))
    with open(out_prefix + ".json", "w") as fh:
        json.dump(recs, fh)


# --------------------------------------------------------------------------- training
def train(X, Y, H, epochs: int = 60, batch: int = 64, lr: float = 1e-3, seed: int = 0, device="cuda", log=print,
          init=None):
    """Adam over the labelled joins; from He-initialised weights, or from a
    copy of `init` (fine-tuning)."""
    torch = _t()
    w = init_weights(seed, device) if init is None else init.detach().clone().float().to(device)
    m = torch.zeros_like(w)
    v = torch.zeros_like(w)
    Xd = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).to(device)
    Yd = torch.from_numpy(np.ascontiguousarray(Y, dtype=np.int32)).to(device)
    Hd = torch.from_numpy(np.ascontiguousarray(H, dtype=np.float32)).to(device)
    n = Xd.shape[0]
    rng = np.random.default_rng(seed)
    step = 0
    for ep in range(epochs):
        order = torch.from_numpy(rng.permutation(n)).to(device)
        tot = 0.0
        for b0 in range(0, n, batch):
            idx = order[b0:b0 + batch]
            step += 1
            loss = train_step(w, m, v, step, lr, Xd[idx], Yd[idx], Hd[idx])
            tot += float(loss.item()) * idx.numel()
        if ep % 10 == 0 or ep == epochs - 1:
            log(f"epoch {ep}: mean loss {tot / n:.4f}")
    return w


def evaluate(w, X, Y, H, C, E, recs, device="cuda"):
    """Plan accuracy, fallback rate and time regret (chosen plan time / best
    plan time) of the planner on labelled joins, and the plain plan's regret
    on the joins where both plans were among the timed candidates."""
    from . import plan_default
    torch = _t()
    logits = forward(w, torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).to(device)).cpu().numpy()
    hits, fallbacks, regrets, paired, wins, losses = 0, 0, [], [], 0, 0
    for i, rec in enumerate(recs):
        sp = rec["spec"]
        plan, fb = decode(logits[i], C[i], E[i], sp["n_build"], sp["n_probe"], sp["key_bytes"])
        fallbacks += int(fb)
        key = f"{plan.passes}:{plan.bits1}:{plan.bits2}"
        best = min(rec["timings"].values())
        t = rec["timings"].get(key)
        if t is not None:
            regrets.append(t / best)
        plain = plan_default(sp["n_build"], sp["n_probe"], sp["key_bytes"])
        tp = rec["timings"].get(f"{plain.passes}:{plain.bits1}:{plain.bits2}")
        if t is not None and tp is not None:
            paired.append((t / best, tp / best))
            wins += int(t < tp)
            losses += int(t > tp)
        hits += int(list(rec["best"]) == [plan.passes, plan.bits1, plan.bits2])
    n = len(recs)
    pr = np.array(paired).reshape(-1, 2)
    return {"joins": n, "plan_accuracy": hits / n, "fallback_rate": fallbacks / n,
            "mean_regret": float(np.mean(regrets)) if regrets else None,
            "timed_choices": len(regrets),
            "paired_joins": len(pr), "paired_mean_regret": float(pr[:, 0].mean()) if len(pr) else None,
            "plain_mean_regret": float(pr[:, 1].mean()) if len(pr) else None,
            "learned_faster": wins, "learned_slower": losses}


def main(argv: Optional[list] = None):
    import argparse
    ap = argparse.ArgumentParser(description="label training joins or train the planner")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("label")
    a.add_argument("--first", type=int, required=True)
    a.add_argument("--count", type=int, required=True)
    a.add_argument("--out", required=True)
    b = sub.add_parser("train")
    b.add_argument("--data", nargs="+", required=True, help=".npz/.json prefixes from `label`")
    b.add_argument("--epochs", type=int, default=60)
    b.add_argument("--holdout", type=float, default=0.2)
    b.add_argument("--weights", default=WEIGHTS_PATH, help="where to write the trained weights")
    b.add_argument("--lr", type=float, default=1e-3)
    b.add_argument("--seed", type=int, default=0)
    b.add_argument("--init", default=None, help="weights to fine-tune instead of a fresh initialisation")
    e = sub.add_parser("evaluate", help="score weights on labelled joins against the plain plan")
    e.add_argument("--data", nargs="+", required=True, help=".npz/.json prefixes from `label`")
    e.add_argument("--holdout", type=float, default=0.2)
    e.add_argument("--weights", default=WEIGHTS_PATH)
    args = ap.parse_args(argv)
    if args.cmd == "label":
        build_dataset(args.first, args.count, args.out, log=lambda s: print(s, flush=True))
        return
    X, Y, H, C, E, recs = [], [], [], [], [], []
    for pre in args.data:
        z = np.load(pre + ".npz")
        X.append(z["X"]); Y.append(z["Y"]); H.append(z["H"]); C.append(z["C"]); E.append(z["E"])
        with open(pre + ".json") as fh:
            recs += json.load(fh)
    X, Y, H, C, E = (np.concatenate(t) for t in (X, Y, H, C, E))
    H = np.stack([heavy_targets(C[i], E[i], recs[i]["spec"]["n_probe"]) for i in range(len(recs))])
    n = len(recs)
    n_test = int(n * args.holdout)
    tr = slice(0, n - n_test)
    te = slice(n - n_test, n)
    if args.cmd == "evaluate":
        w = load_weights(args.weights)
        print(json.dumps({"weights": args.weights,
                          "train": evaluate(w, X[tr], Y[tr], H[tr], C[tr], E[tr], recs[tr]),
                          "holdout": evaluate(w, X[te], Y[te], H[te], C[te], E[te], recs[te])}), flush=True)
        return
    init = load_weights(args.init) if args.init else None
    w = train(X[tr], Y[tr], H[tr], epochs=args.epochs, lr=args.lr, seed=args.seed, init=init,
              log=lambda s: print(s, flush=True))
    ev_tr = evaluate(w, X[tr], Y[tr], H[tr], C[tr], E[tr], recs[tr])
    ev_te = evaluate(w, X[te], Y[te], H[te], C[te], E[te], recs[te])
    print(json.dumps({"train": ev_tr, "holdout": ev_te}), flush=True)
    meta_path = os.path.splitext(args.weights)[0] + ".json"
    save_weights(w, {"features": FEATURES, "hidden": HIDDEN, "logits": LOGITS, "train": ev_tr, "holdout": ev_te,
                     "joins": n, "epochs": args.epochs, "lr": args.lr, "seed": args.seed,
                     "fine_tuned_from": os.path.basename(args.init) if args.init else None,
                     "data": "seeded training joins i = 0..n-1 of planner.join_spec"}, args.weights, meta_path)


if __name__ == "__main__":
    main()
