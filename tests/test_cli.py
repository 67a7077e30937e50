"""The `partitioned_join` command-line driver: the reference binary's flags
(main.cu:100-120) and CSV columns (scripts/bench_partitioned_join.sh:12).

Argument errors are checked on the CPU (they are rejected before any device
call).  On the GPU every generator mode and the binary key-file input run
small joins, and the driver's own CPU check (the reference's check_results,
src/inner_join_cpu.hpp:109-168) must report zero errors in the last CSV
column."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "partitioned_join_amd", "partitioned_join")

# build, probe, key bytes, tuple bytes, selectivity, unique, P1, P2, factor,
# skew mode, zipf, joined, histogram ms, P1 ms, P2 ms, skew ms, join ms, total ms, errors
CSV_FIELDS = 19


def _run(args, timeout=120):
    assert os.path.exists(CLI), "build the driver first (make -C partitioned_join_amd)"
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)


def test_help_and_argument_errors():
    r = _run(["-h"])
    assert r.returncode == 1 and "Usage: partitioned_join" in r.stdout
    r = _run(["-b", "1000", "-n", "3"])
    assert r.returncode == 1 and "power of two" in r.stderr
    r = _run(["-R", "-n", "16", "-N", "12"])
    assert r.returncode == 1 and "power of two" in r.stderr
    r = _run(["-K", "5"])
    assert r.returncode == 1 and "-K must be 4 or 8" in r.stderr
    r = _run(["-b", str(1 << 31)])
    assert r.returncode == 1 and "table sizes" in r.stderr
    r = _run(["--no-such-flag"])
    assert r.returncode == 1


def _csv(args):
    r = _run(["-C", *args])
    assert r.returncode == 0, (args, r.stdout[-2000:], r.stderr[-2000:])
    line = r.stdout.strip().splitlines()[-1].split(",")
    assert len(line) == CSV_FIELDS, line
    return line


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    ["-b", "100000", "-p", "400000"],                                   # one pass, duplicate build keys
    ["-b", "100000", "-p", "400000", "-u", "-s", "0.5"],                # unique build keys, lottery probes
    ["-b", "200000", "-p", "800000", "-R", "-n", "16", "-N", "32"],     # two passes
    ["-b", "100000", "-p", "300000", "-V", "-R", "-n", "8", "-N", "8"],  # payload columns
    ["-b", "50000", "-p", "200000", "-Z", "2", "-z", "1.0"],            # Zipf probes over unique keys
    ["-b", "50000", "-p", "200000", "-Z", "1"],                         # skew mode 1: full selectivity
    ["-b", "100000", "-p", "400000", "-K", "8", "-s", "0.7"],           # 64-bit keys
])
def test_generated_inputs_match_cpu_check(args):
    f = _csv(args)
    assert f[-1] == "0", f
    assert int(f[11]) >= 0 and float(f[17]) > 0


@pytest.mark.gpu
def test_key_files(tmp_path):
    rng = np.random.default_rng(5)
    build = rng.permutation(1 << 17).astype(np.int32)
    probe = rng.integers(0, 1 << 18, 1 << 19).astype(np.int32)
    bf, pf = tmp_path / "build.bin", tmp_path / "probe.bin"
    build.tofile(bf)
    probe.tofile(pf)
    f = _csv(["-F", "-B", str(bf), "-P", str(pf), "-b", str(build.size), "-p", str(probe.size), "-R", "-n", "16",
              "-N", "16"])
    assert f[-1] == "0"
    assert int(f[11]) == int(np.count_nonzero(probe < (1 << 17)))
    # a file shorter than the requested size is an error, not a short read
    r = _run(["-C", "-F", "-B", str(bf), "-P", str(pf), "-b", str(build.size + 1), "-p", str(probe.size)])
    assert r.returncode == 1 and "fewer than" in r.stderr


def _write_tpch(tmp_path, n_orders, rng):
    """Synthetic dbgen-style orders.tbl / lineitem.tbl: sparse unique order
    keys, 1-7 lines per order, priorities '1-URGENT' .. '5-LOW'."""
    prios = ["1-URGENT", "2-HIGH", "3-MEDIUM", "4-NOT SPECIFIED", "5-LOW"]
    okeys = (np.arange(n_orders, dtype=np.int64) // 8) * 32 + (np.arange(n_orders) % 8) + 1
    oprio = rng.integers(0, 5, n_orders)
    with open(tmp_path / "orders.tbl", "w") as f:
        for k, p in zip(okeys.tolist(), oprio.tolist()):
            f.write(f"{k}|{k % 1000}|O|1234.56|1996-01-02|{prios[p]}|Clerk#000000951|0|regular deposits|\n")
    lines = rng.integers(1, 8, n_orders)
    lkeys = np.repeat(okeys, lines)
    with open(tmp_path / "lineitem.tbl", "w") as f:
        for i, k in enumerate(lkeys.tolist()):
            f.write(f"{k}|{i % 977}|{i % 31}|{i % 7 + 1}|17|21168.23|0.04|0.02|N|O|1996-03-13|1996-02-12|"
                    f"1996-03-22|DELIVER IN PERSON|TRUCK|egular courts above the|\n")
    return okeys, oprio, lkeys


@pytest.mark.gpu
def test_tpch_q4_tables(tmp_path):
    """TPC-H Q4 input (-F with .tbl files): build = orders' o_orderkey with the
    priority's first character as payload, probe = lineitem's l_orderkey; a
    probe request past the file's rows repeats the file."""
    rng = np.random.default_rng(11)
    okeys, oprio, lkeys = _write_tpch(tmp_path, 20000, rng)
    n_probe = len(lkeys) + 12345
    f = _csv(["-F", "-B", str(tmp_path / "orders.tbl"), "-P", str(tmp_path / "lineitem.tbl"), "-b", str(len(okeys)),
              "-p", str(n_probe), "-R", "-n", "16", "-N", "16"])
    assert f[-1] == "0", f
    assert int(f[11]) == n_probe  # every lineitem row (and repeat) has exactly one order
    r = _run(["-C", "-F", "-B", str(tmp_path / "lineitem.tbl"), "-P", str(tmp_path / "missing.tbl"), "-b", "10",
              "-p", "10"])
    assert r.returncode == 1 and "cannot open" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("mode,z", [(1, "0.25"), (2, "0.5"), (3, "0.75"), (2, "1"), (3, "1")])
def test_skewed_data_script_flags(mode, z):
    """The flags of every line of the reference's scripts/bench_skewed_data.sh
    (-R -n 512 -N 512 -C -z Z -Z MODE, i.e. 2^18 partitions in two passes)
    with the table sizes scaled from 32M to 1M: the driver's CPU check reports
    zero errors."""
    f = _csv(["-b", "1000000", "-p", "1000000", "-r", "2000000", "-R", "-n", "512", "-N", "512", "-z", z, "-Z",
              str(mode)])
    assert f[6] == "512" and f[7] == "512", f
    assert f[-1] == "0", f
