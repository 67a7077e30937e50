"""Writes the golden vectors under tests/golden/.

  cudf_kat.json      inner-join known answers restated from the cuDF
                     documentation and tests (cudf:include/cudf/join.hpp:73-75,
                     cudf:tests/join/join_tests.cpp:1211-1244,1314-1340), plus
                     the key the reference reserves as its empty-slot marker
                     (-1, src/hash_join_kernels.cuh:20,86,106), which must join
                     like any other key here.

The file is data only: inputs and the expected pairs, written by hand from
the join's definition {(b, p) : R.key[b] = S.key[p]}.  The reference's own
CPU join is not built in this image (its headers need the CUDA toolkit), so
no vector here comes from running it.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
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
    with open(os.path.join(HERE, "cudf_kat.json"), "w") as f:
        json.dump(kat, f, separators=(",", ":"))
    print("wrote cudf_kat.json")


if __name__ == "__main__":
    main()
