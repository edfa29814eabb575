"""Measure the reference traversal's per-query visit counts on the benchmark
workload (2^20-triangle cut torus, queries uniform in the 10%-padded AABB) with
the CPU restatement, and write tests/golden/traversal_counts.json.

These counts define the algorithmic bytes per query used for the roofline in
bench.py: box tests, leaf evaluations and silhouette-edge tests made by the
reference's visit order (query.cuh:79-423)."""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from snch_lbvh_amd import workloads as W  # noqa: E402


def main(sample=4096):
    V, F = W.cut_torus(1024, 512, seed=3)
    lo, hi = W.padded_box(V)
    q = W.uniform_points(1 << 22, lo, hi, seed=4)[:sample]
    d = W.unit_directions(1 << 22, 5)[:sample]
    t0 = time.time()
    tree = oracle.build_tree(V, F)
    print("cpu build", time.time() - t0, "s")
    out = {"workload": "cut torus 1024x512 (2^20 triangles), seed 3; first %d of 2^22 queries (seed 4)" % sample}
    _, _, c = oracle.closest_silhouette(tree, q, counts=True)
    out["silhouette"] = {"box_tests": float(c[0]) / sample, "leaf_evals": float(c[1]) / sample,
                         "edge_tests": float(c[2]) / sample}
    _, _, c = oracle.closest_point(tree, q, counts=True)
    out["closest_point"] = {"box_tests": float(c[0]) / sample, "leaf_evals": float(c[1]) / sample}
    *_, c = oracle.ray(tree, q, d, counts=True)
    out["ray"] = {"box_tests": float(c[0]) / sample, "leaf_evals": float(c[1]) / sample}
    path = os.path.join(ROOT, "tests", "golden", "traversal_counts.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 4096)
