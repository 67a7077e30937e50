"""CPU check of libpj's host code: its sources built for the host only with
AddressSanitizer (tests/host/Makefile), against a stand-in HIP runtime that
runs no kernel but checks every launch -- grid, block, LDS, and for the join
pipeline's kernels that each buffer argument lies inside one allocation over
the extent the kernel can touch.  The driver replays C-ABI call sequences:
eager joins and batched joins created, launched and destroyed in the order
the batch tests and the autotuner use, with their per-join plans."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HOST = os.path.join(HERE, "host")
DRIVER = os.path.join(HOST, "_build", "batch_driver")


@pytest.fixture(scope="module")
def driver():
    subprocess.run(["make", "-C", HOST, "-j8", "-s"], check=True, capture_output=True)
    return DRIVER


def run(driver, script):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0")
    r = subprocess.run([driver], input=script, capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    return r.stdout


def small_batch_sizes(count, seed):
    """The table sizes of tests/test_batch.py::_small_batch (same generator
    calls), plus its empty join."""
    rng = np.random.default_rng(seed)
    sizes = []
    for j in range(count):
        nb = int(rng.integers(1000, 120000))
        ns = int(nb * rng.uniform(1, 6))
        if j % 3 != 0:
            rng.uniform(0, 1.2)
        sizes.append((nb, ns))
    sizes.append((0, 100))
    return sizes


def plan_line(p):
    return (f"{p.passes} {p.bits1} {p.bits2} {p.max_partition_size_factor!r} {p.unique_build_keys} "
            f"{p.probe_slice} {p.source}")


def joins_block(sizes, kb=8):
    return f"joins {len(sizes)}\n" + "".join(f"{nb} {ns} {kb}\n" for nb, ns in sizes)


def test_batch_replay_sequence(driver):
    # tests/test_batch.py::test_batch_graph_replays_match_eager: 24 joins + an empty one, 4 chains
    sizes = small_batch_sizes(24, 9)
    out = run(driver, joins_block(sizes) + "eager\ncreate 0 4 null\nlaunch 0 4\ndestroy 0\n")
    assert "launch errors 0" in out


def test_batch_autotune_sequence_with_planned_joins(pj, driver):
    """The sequence of tests/test_batch.py::test_batch_autotune_and_learned_plans:
    12 int64 joins + an empty one, eager joins, then six batches (chains 2/4 x
    fan-out -1/0/+1) each launched three times and destroyed except the kept
    one, then the kept one launched again.  On these joins the learned planner
    falls back to the plain plan (its heavy-key prediction disagrees with the
    exact counts), so the plans are the plain plans shifted as the autotuner
    shifts them."""
    from partitioned_join_amd import batch
    sizes = small_batch_sizes(12, 10)
    base = [pj.plan_default(nb, ns, 8) for nb, ns in sizes]
    script = joins_block(sizes) + "eager\n"
    bid = 0
    kept = None
    for delta in (-1, 0, 1):
        plans = [batch.shift_plan(p, delta, nb, 8) for p, (nb, _) in zip(base, sizes)]
        for chains in (2, 4):
            script += f"create {bid} {chains} plans\n" + "".join(plan_line(p) + "\n" for p in plans)
            script += f"launch {bid} 3\n"
            # keep every other configuration as the "fastest" so far, as timings would
            if kept is None or bid % 2 == 1:
                if kept is not None:
                    script += f"destroy {kept}\n"
                kept = bid
            else:
                script += f"destroy {bid}\n"
            bid += 1
    script += f"launch {kept} 1\ndestroy {kept}\n"
    out = run(driver, script)
    assert "launch errors 0" in out and "call failures 0" in out


@pytest.mark.parametrize("kb", [4, 8])
def test_every_plan_shape_on_small_and_ragged_joins(pj, driver, kb):
    """Every pass count and radix-bit split up to 24 bits (above 16 the pass-2
    digits are counted after pass 1; every split with 11 or 12 pass-1 bits,
    whose pass-2 item prefix takes the most LDS) on tables from 1 row to 300K
    rows: the host carve-up and every launch stay inside the buffers and the
    LDS."""
    sizes = [(1, 1), (1, 5000), (63, 64), (1000, 7), (4097, 20000), (100000, 300000), (0, 10), (10, 0)]
    cap = pj.lib().pj_join_chunk_capacity(kb)
    script = joins_block(sizes, kb)
    bid = 0
    shapes = [(1, b, 0) for b in range(1, 13)] + [(2, b1, b2) for b1 in range(1, 13) for b2 in range(1, 13)
                                                  if b1 + b2 <= 16 or (b1 + b2) % 3 == 0 or b1 >= 11]
    for passes, b1, b2 in shapes:
        ok = all(1.5 * nb / 2 ** (b1 + b2) <= cap for nb, _ in sizes)
        if not ok:
            continue
        lines = "".join(f"{passes} {b1} {b2} 1.5 0 0 2\n" for _ in sizes)
        script += f"create {bid} 3 plans\n{lines}launch {bid} 1\ndestroy {bid}\n"
        bid += 1
    out = run(driver, script)
    assert "launch errors 0" in out and "call failures 0" in out


def test_checks_catch_a_short_workspace(driver):
    # planted mistake: the workspace the stand-in runtime knows is 4 KB shorter
    # than the size the join was given, so launches that use its tail overrun
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([driver], input=joins_block([(100000, 300000)]) + "eager_short\n", capture_output=True,
                       text=True, env=env, timeout=120)
    assert r.returncode == 1 and "overruns eager workspace (short)" in r.stderr, r.stderr[-2000:]


def test_noalloc_entry_point_stays_in_its_scratch(driver):
    # the static scratch is registered with the stand-in runtime at its declared size;
    # joins from 1 row to the largest that fits (and one that does not: rejected)
    sizes = [(1, 1), (1000, 7), (1 << 20, 1 << 22), (4 << 20, 8 << 20), (16 << 20, 64 << 20), (0, 5)]
    out = run(driver, joins_block(sizes, 4) + "noalloc\n" + joins_block(sizes, 8) + "noalloc\n")
    assert "launch errors 0" in out and "call failures 0" in out
