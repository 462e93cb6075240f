"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

  c1_pairs.npy    sorted (lo, hi) uint32 pair set of the 16,384-box uniform
                  instance (centres U[0,1024)^2, sides U[1,8], seed 1),
                  cross-checked here by brute force and sweep-and-prune
  manifest.json   per named instance: n, range, reference cell size, pair
                  count, SHA-256 of the sorted pair set, grid statistics at
                  the reference cell size; counts of the reference's own
                  logged 3-D answers (benchmarks/test1/filter_file:7,
                  benchmarks/test3/filter_file:17-49)

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from pbig_amd import workloads  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def entry(name, boxes, L, independent):
    cs = oracle.decide_cell_size(boxes, L)
    pairs, st = oracle.grid_pairs(boxes, L, cs, with_stats=True)
    pairs = oracle.sort_pairs(pairs)
    ind = oracle.sort_pairs(independent(boxes))
    assert np.array_equal(pairs, ind), name
    return pairs, {
        "n": int(boxes.shape[0]), "range": float(L), "reference_cell_size": cs,
        "pairs": int(len(pairs)), "sha256": oracle.pair_hash(pairs),
        "first_box": [float(v) for v in boxes[0]],
        "grid_stats_at_reference_cell_size": st,
    }


def main():
    oracle.build()
    man = {"generator": "pbig_amd.workloads (SplitMix64, deterministic math)"}
    b1, L1 = workloads.c1_small_uniform()
    p1, man["c1_uniform_16k"] = entry("c1", b1, L1, oracle.brute_pairs)
    np.save(os.path.join(HERE, "c1_pairs.npy"), p1)
    b2, L2 = workloads.c2_large_uniform()
    _, man["c2_uniform_4m"] = entry("c2", b2, L2, oracle.sweep_pairs)
    b3, L3 = workloads.c3_clustered()
    _, man["c3_clustered_2m"] = entry("c3", b3, L3, oracle.sweep_pairs)
    man["reference_logged_3d"] = {
        "test1_UU_1e7": {"pairs": 49514252, "source": "benchmarks/test1/filter_file:7"},
        "test3_UN_1e7": {"pairs": 1121815231, "cell_size": 128, "n_cells": 493039,
                         "memberships": 26668901, "max_per_cell": 3543, "max_cell": 246442,
                         "jobs256": 200096, "compared_pairs": 16683783392,
                         "source": "benchmarks/test3/filter_file:5-49"},
    }
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(man, f, indent=1, sort_keys=True)
    print(json.dumps(man, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
