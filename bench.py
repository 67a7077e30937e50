"""Benchmark: radix-partitioned join of 64M unique int32 build keys with 256M
int32 probe keys (50 % match), row-id payloads, two radix passes, on one
MI355X.  One step = one full join (histogram, both partition passes, work list,
LDS hash join, output compaction) with inputs resident in HBM.

Prints one JSON line: whole-job tuples/s, the roofline of the dominant kernel
(algorithmic bytes / its hipEvent-timed duration vs 8 TB/s), per-phase times,
the plain and learned plans on the Zipf(1.0) skewed join and on eight seeded
joins the planner was not trained on, the batched-joins record, and a CPU
baseline: the oracle's restatement of the reference's
inner_join_cpu (std::unordered_map, serial build, OpenMP probe; the
reference's own CPU code needs the CUDA toolkit and is not built here) timed
on this host's cores on the full 64M x 256M workload, the 1M x 4M one and the
16M x 256M Zipf(1.0) one.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "tuples/s & % HBM roofline, 64M⋈256M int32; learned-plan vs plain at Zipf 1.0"


def algorithmic_bytes(n_r, n_s, n_out, passes, w=8):
    """Per-phase algorithmic bytes: each pass reads and writes every tuple once,
    the join reads every tuple once and writes 8 bytes per output pair, the
    histogram reads every key once (reported, not part of the headline total)."""
    t = n_r + n_s
    return {
        "histogram": t * 4,
        "pass1": 2 * t * w,
        "pass2": 2 * t * w if passes == 2 else 0,
        "join": t * w + 8 * n_out,
        "total": passes * 2 * t * w + t * w + 8 * n_out,
    }


PHASES = ["histogram_ms", "offsets_ms", "pass1_ms", "pass2_ms", "skew_ms", "join_ms", "compact_ms"]


def run_join_steps(pj, torch, bk, pk, plan, steps, warmup, bp=None, pp=None, n_cap=None, marks=0xFF):
    """warmup + timed steps of the async join, queued back to back; returns
    (ms/step, phase means, count).  The phase means come from the HIP events
    each timed join records (the last 128 when there are more), read after
    the timed region; `marks` selects the phase boundaries recorded (phases
    without both of theirs read 0)."""
    n_cap = n_cap if n_cap is not None else pk.numel() + 1024
    out = torch.empty((n_cap, 2), dtype=torch.int32, device="cuda")
    ws_bytes = pj.workspace_size(bk.numel(), pk.numel(), bk.element_size(), plan)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    for _ in range(warmup):
        d_count = pj.partitioned_join_async(bk, pk, out, bp, pp, plan=plan, workspace=ws)
    torch.cuda.synchronize()
    pj.profile_enable(True)
    pj.profile_select(marks)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        d_count = pj.partitioned_join_async(bk, pk, out, bp, pp, plan=plan, workspace=ws)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phases = pj.profile_read_recent(steps)
    assert len(phases) == min(steps, 128) and all(p["valid"] for p in phases), len(phases)
    pj.profile_select(0xFF)
    pj.profile_enable(False)
    count = int(d_count.item())
    keys = [k for k in phases[0] if k.endswith("_ms")]
    mean = {k: sum(p[k] for p in phases) / len(phases) for k in keys}
    return (t1 - t0) * 1000.0 / steps, mean, count


def copy_calibration(torch, gib=1, reps=3):
    """A plain device copy of a known size, run under the same rocprofv3 --pmc
    pass as the join: its FETCH_SIZE / WRITE_SIZE per launch against the exact
    bytes moved give the factor by which the counters are scaled."""
    n = (gib << 30) // 4
    src = torch.ones(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    torch.cuda.empty_cache()


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (profiles/traffic_64Mx256M.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_64Mx256M.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get("kernels", {}).get(kernel)
    return None if rec is None else rec.get("bytes_per_launch")


def batched_joins(pj, torch, count=512):
    """512 independent int64-key joins of uneven sizes replayed as one HIP
    graph: the untuned default (8 chains, plain plans) and the autotuned
    configuration over chains x fan-out, with plain per-join plans."""
    from partitioned_join_amd import batch
    builds, probes, specs = batch.make_batch_workload(count)
    tuples = sum(s["n_build"] + s["n_probe"] for s in specs)
    plain = batch.batch_plans(builds, probes, "plain")
    bj = batch.BatchJoin(builds, probes, plain, chains=8)
    ms_default = bj.time(3)
    counts_default = bj.counts()
    bj.close()
    t0 = time.perf_counter()
    best, cfg, timings = batch.autotune(builds, probes, plain, chain_options=(8, 16, 32), bit_offsets=(-1, 0, 1),
                                        tiles=((None, None), (1.0, 8192)), reps=3)
    tune_s = time.perf_counter() - t0
    ms_best = best.time(5)
    counts_best = best.counts()
    best.close()
    assert np.array_equal(counts_default, counts_best)
    # the same joins one after another on one stream (no graph, plain plans,
    # one shared workspace): what batching is measured against
    ws = torch.empty(max(pj.workspace_size(b.numel(), p.numel(), 8, pl) for b, p, pl in zip(builds, probes, plain)),
                     dtype=torch.uint8, device="cuda")
    outs = [torch.empty((max(1, p.numel()), 2), dtype=torch.int32, device="cuda") for p in probes]
    def sequential():
        return [pj.partitioned_join_async(b, p, o, plan=pl, workspace=ws) for b, p, o, pl in zip(builds, probes, outs,
                                                                                                 plain)]
    sequential()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        seq_counts = sequential()
    torch.cuda.synchronize()
    ms_seq = (time.perf_counter() - t0) * 1000.0 / 3
    assert np.array_equal(np.array([int(c.item()) for c in seq_counts]), counts_best)
    del ws, outs
    return {"workload": f"{count} joins, |R| log-uniform [10K,4M], |S| 1-8x|R|, uniform 64-bit int64 build keys, "
                        "Zipf(theta~U[0,1.2]) probes, int32 row-id payloads, one HIP graph",
            "tuples": tuples, "pairs": int(counts_best.sum()),
            "default_ms": ms_default, "default_tuples_per_s": tuples / (ms_default / 1000.0),
            "tuned_ms": ms_best, "tuned_tuples_per_s": tuples / (ms_best / 1000.0),
            "sequential_ms": ms_seq, "sequential_tuples_per_s": tuples / (ms_seq / 1000.0),
            "tuned_config": {"chains": cfg[0], "fanout_offset_bits": cfg[1],
                             "join_tile": "plan's own" if len(cfg) == 2 else {"chunk_factor": cfg[2][0],
                                                                               "probe_slice": cfg[2][1]},
                             "plans": "plain"},
            "autotune_seconds": tune_s,
            "timings_ms": {",".join(str(x) for x in k): v for k, v in timings.items()}}


def batched_joins_child(timeout_s=600):
    """The batched-joins record, measured in a child process: if that run
    fails, the record says so and the main measurement is still printed."""
    import subprocess
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--batch-only"], capture_output=True,
                           text=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return {"error": f"batched-joins run exceeded {timeout_s} s"}
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"batched-joins run ended with status {r.returncode}: {r.stderr[-500:]}"}
    return json.loads(lines[-1])


def variant_records(pj, torch, steps):
    """The same join shape with 64-bit keys, with explicit payload columns and
    with duplicate build keys: tuples/s and the whole-job roofline fraction
    with w = key + payload bytes."""
    G = pj.generators
    recs = {}
    n_r, n_s = 1 << 26, 1 << 28
    L = pj.lib()
    s = torch.cuda.current_stream().cuda_stream
    # 64-bit keys: uniform random build keys (unique with overwhelming probability), half the probes match
    bk = torch.empty(n_r, dtype=torch.int64, device="cuda")
    pk = torch.empty(n_s, dtype=torch.int64, device="cuda")
    pj._lib.check(L.pj_gen_build_i64(bk.data_ptr(), n_r, 0x64B, s))
    pj._lib.check(L.pj_gen_probe_i64(pk.data_ptr(), n_s, n_r, 0x64B, 0x64F, 0.5, s))
    plan = pj.plan_default(n_r, n_s, 8)
    ms, ph, count = run_join_steps(pj, torch, bk, pk, plan, steps, 1)
    ab = algorithmic_bytes(n_r, n_s, count, plan.passes, w=12)
    recs["int64_keys"] = {"workload": "64M uniform 64-bit build keys x 256M probes, 50% match, row-id payloads",
                          "ms": ms, "tuples_per_s": (n_r + n_s) / (ms / 1000.0), "pairs": count,
                          "whole_job_roofline_frac": ab["total"] / (ms / 1000.0) / 1e9 / HBM_PEAK_GBS,
                          "plan": plan.as_dict(), "phases_ms": ph}
    log(f"int64 64M x 256M: {ms:.3f} ms")
    del bk, pk
    torch.cuda.empty_cache()
    # payload columns: the 64M x 256M workload with random int32 payloads on both sides
    bk, pk = G.workload_config2()
    g = torch.Generator(device="cuda").manual_seed(0xBA1)
    bp = torch.randint(-2**31, 2**31 - 1, (n_r,), generator=g, device="cuda", dtype=torch.int32)
    pp = torch.randint(-2**31, 2**31 - 1, (n_s,), generator=g, device="cuda", dtype=torch.int32)
    plan = pj.plan_default(n_r, n_s, 4)
    plan.unique_build_keys = 1  # the reference's payload benchmark passes -u -V
    ms, ph, count = run_join_steps(pj, torch, bk, pk, plan, steps, 1, bp, pp)
    ab = algorithmic_bytes(n_r, n_s, count, plan.passes, w=8)
    recs["payload_columns"] = {"workload": "64M x 256M int32 (the headline keys) with int32 payload columns on both "
                                           "sides", "ms": ms, "tuples_per_s": (n_r + n_s) / (ms / 1000.0),
                               "pairs": count,
                               "whole_job_roofline_frac": ab["total"] / (ms / 1000.0) / 1e9 / HBM_PEAK_GBS,
                               "phases_ms": ph}
    log(f"payload columns 64M x 256M: {ms:.3f} ms")
    del bp, pp, pk
    # duplicate build keys: every value of [0, 2^24) four times; probes uniform over [0, 2^25)
    bk = torch.div(bk, 4, rounding_mode="floor")
    pk = torch.randint(0, 1 << 25, (1 << 26,), generator=g, device="cuda", dtype=torch.int32)
    plan = pj.plan_default(bk.numel(), pk.numel(), 4)
    ms, ph, count = run_join_steps(pj, torch, bk, pk, plan, steps, 1, n_cap=4 * pk.numel() + 1024)
    t = bk.numel() + pk.numel()
    ab = algorithmic_bytes(bk.numel(), pk.numel(), count, plan.passes, w=8)
    recs["duplicate_build_keys"] = {"workload": "64M build keys (each of 2^24 values 4 times) x 64M probes "
                                                "uniform over [0, 2^25)", "ms": ms, "tuples_per_s": t / (ms / 1000.0),
                                    "pairs": count, "pairs_per_s": count / (ms / 1000.0),
                                    "whole_job_roofline_frac": ab["total"] / (ms / 1000.0) / 1e9 / HBM_PEAK_GBS,
                                    "phases_ms": ph}
    log(f"duplicate keys 64M x 64M: {ms:.3f} ms")
    del bk, pk
    torch.cuda.empty_cache()
    return recs


def learned_held_out(pj, torch, planner, w, joins=8, first=2000):
    """The learned planner on joins it was not trained on: the first `joins`
    seeded joins from index `first` on (training used indices below it) with at
    most 16M build and 128M probe rows.  Each is timed with the plain plan and,
    when it differs, with the learned plan (same pair count asserted); the
    summary is the geometric mean of plain / learned time with the plan reused
    and with planning included."""
    rows = []
    i = first
    while len(rows) < joins:
        sp = planner.join_spec(i)
        i += 1
        if sp["n_build"] > (16 << 20) or sp["n_probe"] > (128 << 20):
            continue
        bk, pk = planner.make_join(sp)
        torch.cuda.synchronize()
        plain = pj.plan_default(bk.numel(), pk.numel(), bk.element_size())
        planner.plan_learned(bk, pk, w)
        t0 = time.perf_counter()
        lp, fb = planner.plan_learned(bk, pk, w)
        plan_ms = (time.perf_counter() - t0) * 1000.0
        cap = pk.numel() + 1024
        pms, _, pc = run_join_steps(pj, torch, bk, pk, plain, 3, 1, n_cap=cap)
        same = (lp.passes, lp.bits1, lp.bits2) == (plain.passes, plain.bits1, plain.bits2)
        lms = pms
        if not same:
            lms, _, lc = run_join_steps(pj, torch, bk, pk, lp, 3, 1, n_cap=cap)
            assert lc == pc, (lc, pc)
        rows.append({"join": sp["i"], "build_rows": sp["n_build"], "probe_rows": sp["n_probe"],
                     "key_bytes": sp["key_bytes"], "zipf_theta": round(sp["theta"], 3),
                     "match": round(sp["match"], 3), "plain_plan": [plain.passes, plain.bits1, plain.bits2],
                     "learned_plan": [lp.passes, lp.bits1, lp.bits2], "fallback_to_plain": fb,
                     "plain_ms": pms, "learned_ms": lms, "planning_ms": plan_ms})
        del bk, pk
        torch.cuda.empty_cache()
    reuse = [r["plain_ms"] / r["learned_ms"] for r in rows]
    incl = [r["plain_ms"] / (r["learned_ms"] + r["planning_ms"]) for r in rows]
    gm = lambda v: math.exp(sum(math.log(x) for x in v) / len(v))
    return {"workload": f"seeded held-out joins {first}.. (build <= 16M, probe <= 128M rows, int32 or int64 keys, "
                        "Zipf probes with misses)", "joins": rows,
            "plans_differing": sum(r["plain_plan"] != r["learned_plan"] for r in rows),
            "learned_faster": sum(r["learned_ms"] < r["plain_ms"] for r in rows),
            "learned_slower": sum(r["learned_ms"] > r["plain_ms"] for r in rows),
            "speedup_plan_reused_geomean": gm(reuse), "speedup_planning_included_geomean": gm(incl)}


def log(msg):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for l in f:
                if l.startswith("model name"):
                    model = l.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def cpu_join(bk, pk, threads):
    """The oracle's inner_join_cpu (unique build keys) on host key arrays:
    (tuples/s, pairs, seconds in the omp_get_wtime window of build + probe,
    src/inner_join_cpu.hpp:41-78)."""
    from oracle import oracle as O
    cap = len(pk)  # unique build keys: at most one pair per probe row
    _, n, secs = O.inner_join_cpu(bk, pk, uniq_keys=True, num_threads=threads, capacity=cap, return_seconds=True)
    return (len(bk) + len(pk)) / secs, n, secs


def cpu_baseline(c2_host, c1_host, c3_host, threads):
    """CPU join on the headline workload (64M x 256M, the same keys the GPU
    joined), plus the 1M x 4M and 16M x 256M Zipf(1.0) workloads."""
    v, n, secs = cpu_join(*c2_host, threads)
    rec = {"value": v, "unit": "tuples/s", "cores": threads, "kind": "port",
           "sample": f"full 64M x 256M workload (the bench's own keys, copied to the host): {n} pairs, "
                     f"build+probe {secs:.2f} s in the omp_get_wtime window of inner_join_cpu",
           "host": cpu_info()}
    v1, n1, s1 = cpu_join(*c1_host, threads)
    rec["config1_1Mx4M"] = {"tuples_per_s": v1, "pairs": n1, "seconds": s1}
    if c3_host is not None:
        v3, n3, s3 = cpu_join(*c3_host, threads)
        rec["config3_zipf1_16Mx256M"] = {"tuples_per_s": v3, "pairs": n3, "seconds": s3}
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-skew", action="store_true", help="skip the Zipf(1.0) comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-joins graph measurement")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the 64-bit-key, payload-column and duplicate-key records")
    ap.add_argument("--batch-only", action="store_true",
                    help="print only the batched-joins record (the child process of a full run)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device copy (counter calibration for rocprofv3 --pmc runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards by top radix "
              "digit across GPUs by design, see the sharding section of the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    import partitioned_join_amd as pj
    assert torch.cuda.is_available(), "needs an MI355X"
    if args.batch_only:
        print(json.dumps(batched_joins(pj, torch)), flush=True)
        return
    G = pj.generators
    bk, pk = G.workload_config2()
    torch.cuda.synchronize()
    plan = pj.plan_default(bk.numel(), pk.numel())
    # The build keys are unique, and the plan says so, as the reference's own
    # benchmark does (-u, scripts/bench_partitioned_join.sh:19): a probe stops
    # at its first match and the build skips its duplicate check.
    plan.unique_build_keys = 1
    # Every phase's events first (untimed: the phase breakdown and the
    # dominant kernel), then the timed steps with the two events around the
    # dominant kernel only: each event costs the GPU a few microseconds (all
    # eight: 0.027 ms of a 3.4 ms join).
    ms_all, ph, count = run_join_steps(pj, torch, bk, pk, plan, args.steps, args.warmup)
    kern = {"pass1": ph["pass1_ms"], "pass2": ph["pass2_ms"], "join": ph["join_ms"],
            "histogram": ph["histogram_ms"]}
    dom = max(kern, key=kern.get)
    i_dom = PHASES.index(dom + "_ms")
    ms, ph_t, count_t = run_join_steps(pj, torch, bk, pk, plan, args.steps, 1, marks=(3 << i_dom))
    assert count_t == count, (count_t, count)
    kern[dom] = ph_t[dom + "_ms"]
    log(f"64M x 256M join: {ms:.3f} ms per step ({ms_all:.3f} with every phase event), {count} pairs")
    if not args.no_variants:  # the same join without the promise: the duplicate check runs
        plan_nu = pj.plan_default(bk.numel(), pk.numel())
        ms_nu, ph_nu, count_nu = run_join_steps(pj, torch, bk, pk, plan_nu, max(3, args.steps // 2), 1)
        assert count_nu == count, (count_nu, count)
    n_r, n_s = bk.numel(), pk.numel()
    ab = algorithmic_bytes(n_r, n_s, count, plan.passes)
    achieved = ab[dom] / (kern[dom] / 1000.0) / 1e9
    value = (n_r + n_s) / (ms / 1000.0)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "radix join, 64M unique int32 build keys (perm of [0,2^26)) x 256M int32 probe "
                               "keys, 50% match (rest in [2^26,2^27)), row-id payloads, output (build_idx, "
                               "probe_idx) pairs; the plan declares the build keys unique (the reference "
                               "benchmark's -u)", "build_rows": n_r, "probe_rows": n_s, "pairs": count,
                   "plan": plan.as_dict()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom), "kernel": dom,
                     "kernel_ms": kern[dom], "algorithmic_bytes": ab[dom]},
        "phases_ms": ph,
        "ms_per_step_every_phase_event": ms_all,
        "whole_job_roofline_frac": ab["total"] / (ms / 1000.0) / 1e9 / HBM_PEAK_GBS,
        "algorithmic_bytes_total": ab["total"],
    }
    if not args.no_variants:
        line["without_unique_promise"] = {"ms": ms_nu, "join_ms": ph_nu["join_ms"], "plan": plan_nu.as_dict()}
    c2_host = (bk.cpu().numpy(), pk.cpu().numpy()) if not args.no_cpu else None
    del bk, pk
    torch.cuda.empty_cache()
    if args.copy_calibration:
        copy_calibration(torch)
    if not args.no_skew:
        zb, zp = G.workload_config3(1.0)
        zplan = pj.plan_default(zb.numel(), zp.numel())
        zplan.unique_build_keys = 1  # unique build keys, as the reference sets for its probe-skew mode (main.cu:346)
        zms, zph, zcount = run_join_steps(pj, torch, zb, zp, zplan, max(3, args.steps // 2), 1)
        rec = {"workload": "16M unique int32 build x 256M Zipf(1.0) probes", "pairs": zcount,
               "plain_tuples_per_s": (zb.numel() + zp.numel()) / (zms / 1000.0), "plain_ms": zms,
               "plain_plan": zplan.as_dict(), "plain_phases_ms": zph, "learned_tuples_per_s": None}
        from partitioned_join_amd import planner
        w = planner.load_weights()
        if w is not None:
            planner.plan_learned(zb, zp, w)  # first call: loads the planner's kernels
            t0 = time.perf_counter()
            for _ in range(3):
                lplan, fallback = planner.plan_learned(zb, zp, w)
            plan_ms = (time.perf_counter() - t0) * 1000.0 / 3
            lplan.unique_build_keys = 1
            lms, lph, lcount = run_join_steps(pj, torch, zb, zp, lplan, max(3, args.steps // 2), 1)
            assert lcount == zcount, (lcount, zcount)
            zt = zb.numel() + zp.numel()
            rec.update({"learned_tuples_per_s": zt / ((lms + plan_ms) / 1000.0),
                        "learned_ms_including_planning": lms + plan_ms,
                        "learned_plan_reuse_tuples_per_s": zt / (lms / 1000.0), "learned_ms": lms,
                        "learned_plan": lplan.as_dict(), "learned_fallback_to_plain": fallback,
                        "learned_planning_ms": plan_ms, "learned_phases_ms": lph,
                        "learned_plan_equals_plain": [lplan.passes, lplan.bits1, lplan.bits2] ==
                                                     [zplan.passes, zplan.bits1, zplan.bits2],
                        "learned_pair_count_equal_plain": lcount == zcount})
            line["learned_held_out"] = learned_held_out(pj, torch, planner, w)
            log("learned planner on held-out joins done")
        line["zipf_theta1"] = rec
        log(f"Zipf(1.0) join: plain {zms:.3f} ms")
        c3_host = (zb.cpu().numpy(), zp.cpu().numpy()) if not args.no_cpu else None
        del zb, zp
        torch.cuda.empty_cache()
    if not args.no_variants:
        line["variants"] = variant_records(pj, torch, max(3, args.steps // 2))
    if not args.no_batch:
        line["batched"] = batched_joins_child()
        log("batched joins done")
    if not args.no_cpu:
        log("CPU baseline (inner_join_cpu on the host cores) ...")
        c1b, c1p = G.workload_config1()
        c1_host = (c1b.cpu().numpy(), c1p.cpu().numpy())
        del c1b, c1p
        line["cpu_baseline"] = cpu_baseline(c2_host, c1_host, None if args.no_skew else c3_host,
                                            min(16, os.cpu_count() or 1))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
