"""Writes the golden vectors under tests/golden/.

Run in a checkout where the reference sources are present and
`make -C oracle ref` has built oracle/_ref/libpjref.so (the reference's own
src/hash_functions.cuh, src/int_fastdiv.h and src/inner_join_cpu.hpp compiled
unchanged).  The vectors are data only: inputs and the reference code's outputs.

  ref_murmur3.json   MurmurHash3_32<int> / <long long> of fixed keys
                     (src/hash_functions.cuh:33-125)
  ref_fastdiv.json   int_fastdiv n / d and n % d (src/int_fastdiv.h:102-123)
  ref_join.json      inner_join_cpu (src/inner_join_cpu.hpp:23-107) outputs
                     {x = key, y = 0} on small seeded tables, sorted
  cudf_kat.json      inner-join known answers restated from the cuDF
                     documentation and tests (cudf:include/cudf/join.hpp:73-75,
                     cudf:tests/join/join_tests.cpp:1211-1244,1314-1340)
"""
import json
import os
import random
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def main():
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libpjref.so not built (make -C oracle ref)")
    rng = random.Random(20261202)
    keys4 = [0, 1, -1, 2, 42, 0x7FFFFFFF, -0x80000000, 123456789, -987654321] + [
        rng.randint(-2**31, 2**31 - 1) for _ in range(247)]
    keys8 = [0, 1, -1, 2**32, 2**32 + 1, 2**63 - 1, -2**63, 0x0123456789ABCDEF] + [
        rng.randint(-2**63, 2**63 - 1) for _ in range(248)]
    murmur = {
        "key4": keys4, "hash4": [R.pjref_murmur3_i32(k) for k in keys4],
        "key8": keys8, "hash8": [R.pjref_murmur3_i64(k) for k in keys8],
    }
    divisors = [1, -1, 2, -2, 3, 7, -7, 10, 641, 2048, 6144, 6145, 3072, 65537, 1 << 30, -(1 << 30), 2**31 - 1,
                -2**31 + 1] + [rng.randint(2, 100000) for _ in range(14)]
    nums = [0, 1, -1, 2, -2, 2**31 - 1, -2**31, 12345, -12345] + [rng.randint(-2**31, 2**31 - 1) for _ in range(55)]
    fastdiv = {"d": divisors, "n": nums,
               "q": [[R.pjref_fastdiv_div(n, d) for n in nums] for d in divisors],
               "r": [[R.pjref_fastdiv_mod(n, d) for n in nums] for d in divisors]}
    joins = []
    for case, (nb, npr, dom, uniq, seed) in enumerate([(1000, 4000, 1000, 1, 11), (4096, 16384, 8192, 1, 12),
                                                       (3000, 5000, 1500, 0, 13), (1, 100, 4, 0, 14)]):
        r = np.random.default_rng(seed)
        if uniq:
            b = r.permutation(dom)[:nb].astype(np.int32)
        else:
            b = r.integers(0, dom, nb).astype(np.int32)
        p = r.integers(0, dom + dom // 2, npr).astype(np.int32)
        xy, n, _ = O.ref_inner_join_cpu(b, p, uniq_keys=uniq)
        xy = xy[np.lexsort((xy[:, 1], xy[:, 0]))]
        joins.append({"build": b.tolist(), "probe": p.tolist(), "uniq": uniq, "count": int(n),
                      "xy": xy.tolist()})
    kat = [
        {"name": "join.hpp example", "build": [0, 1, 2], "probe": [1, 2, 3],
         "pairs": [[1, 0], [2, 1]]},
        {"name": "InnerJoinCornerCase", "build": [2], "probe": [4, 1, 3, 2, 2, 2, 2],
         "pairs": [[0, 3], [0, 4], [0, 5], [0, 6]]},
        {"name": "EqualValuesInnerJoin", "build": [0, 0], "probe": [0, 0],
         "pairs": [[0, 0], [0, 1], [1, 0], [1, 1]]},
        {"name": "empty build", "build": [], "probe": [1, 2], "pairs": []},
        {"name": "empty probe", "build": [1, 2], "probe": [], "pairs": []},
        {"name": "reference sentinel key -1", "build": [-1, 5, -1], "probe": [-1, 7, 5],
         "pairs": [[0, 0], [2, 0], [1, 2]]},
    ]
    for name, obj in [("ref_murmur3.json", murmur), ("ref_fastdiv.json", fastdiv), ("ref_join.json", joins),
                      ("cudf_kat.json", kat)]:
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(obj, f, separators=(",", ":"))
        print("wrote", name)


if __name__ == "__main__":
    main()
