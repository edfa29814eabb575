#!/usr/bin/env python3
"""Benchmark: closest-silhouette queries/s on a 2^20-triangle cut torus with
2^22 queries (plain and learned-bound), plus LBVH builds/s, on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

One step = one closest-silhouette launch over all 2^22 queries with the tree,
the queries and the outputs resident in HBM.  Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per query of the reference's traversal (query.cuh:325-423):
# in 13 B (point + flip) + out 8 B (distance + edge id) + C_box * 52 B (child
# AABB 24 + half a node record 8 + cone 20) + C_leaf * 40 B (triangle record)
# + C_edge * 80 B (silhouette edge 32 + its 4 vertices' 48).  C_box, C_leaf and
# C_edge are the reference visit counts measured by the CPU restatement on this
# workload (oracle/traversal_counts.py -> tests/golden/traversal_counts.json).
S_BOX, S_LEAF, S_EDGE, S_IN, S_OUT = 52, 40, 80, 13, 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--queries", type=int, default=1 << 22)
    ap.add_argument("--nu", type=int, default=1024)
    ap.add_argument("--nv", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1 << 13)
    ap.add_argument("--build-reps", type=int, default=5)
    return ap.parse_args()


def counts_fixture():
    p = os.path.join(ROOT, "tests", "golden", "traversal_counts.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def cpu_baseline(V, F, q, sample, silhouette_ms_gpu):
    """The CPU oracle (plain-C restatement, OpenMP) on a bounded sample of the
    same queries, on this host's cores."""
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    t0 = time.perf_counter()
    tree = oracle.build_tree(V, F)
    t_build = time.perf_counter() - t0
    qs = q[:sample]
    oracle.closest_silhouette(tree, qs[:256], nthreads=threads)  # warm-up
    t0 = time.perf_counter()
    oracle.closest_silhouette(tree, qs, nthreads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": sample / dt,
        "unit": "queries/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample} of the {len(q)} closest-silhouette queries (same seeds), oracle/snch_oracle.c "
                  f"with OpenMP on {threads} threads; CPU build of the same tree took {t_build:.2f} s",
        "build_s": t_build,
    }


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(queries shard without communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch

    import snch_lbvh_amd as S
    from snch_lbvh_amd import workloads as W

    dev = torch.device("cuda:0")
    V, F = W.cut_torus(args.nu, args.nv, seed=3)
    lo, hi = W.padded_box(V)
    q_np = W.uniform_points(args.queries, lo, hi, seed=4)
    t0 = time.perf_counter()
    sc = S.Scene(V, F)
    t_scene = time.perf_counter() - t0
    # builds/s: GPU time of construct() (leaf-gen .. refit), HIP events
    build_ms = []
    for _ in range(args.build_reps + 1):
        sc.build_bvh()
        build_ms.append(sc.build_ms())
    build_ms = float(np.median(build_ms[1:]))

    q = torch.from_numpy(q_np).to(dev)
    n = q.shape[0]
    out_e = torch.empty(n, dtype=torch.int32, device=dev)
    out_d = torch.empty(n, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    def step():
        S.check(S.lib().snch_closest_silhouette(sc._h, S._ptr(q), n, 0, None, S._ptr(out_e), S._ptr(out_d),
                                                S._stream()))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps
    ms_per_step = wall * 1000.0 / args.steps
    value = n * args.steps / wall

    # other queries (reported beside the headline)
    extra = {}

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - a) / reps

    d = torch.from_numpy(W.unit_directions(n, 5)).to(dev)
    r = torch.from_numpy(W.log_uniform(n, 1e-3, 0.5, 6)).to(dev)
    u = torch.rand((n, 3), generator=torch.Generator(device=dev).manual_seed(7), device=dev)
    extra["closest_point_qps"] = n / timed(lambda: sc.closest_point(q))
    extra["ray_qps"] = n / timed(lambda: sc.ray_intersect(q, d))
    extra["sample_in_sphere_qps"] = n / timed(lambda: sc.sample_in_sphere(q, r, u))

    cf = counts_fixture()
    roofline = None
    if cf is not None:
        c = cf["silhouette"]
        per_q = S_IN + S_OUT + c["box_tests"] * S_BOX + c["leaf_evals"] * S_LEAF + c["edge_tests"] * S_EDGE
        achieved = per_q * n / (kernel_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                    "bytes_per_query": per_q, "kernel": "k_closest_silhouette<256>", "kernel_ms": kernel_ms}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(V, F, q_np, args.cpu_sample, kernel_ms)

    line = {
        "metric": "closest-silhouette queries/sec @1M tris (plain vs learned-bound); LBVH builds/sec",
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded cut torus, Perlin displacement; uniform queries)",
        "config": {"workload": f"closest silhouette, cut torus {F.shape[0]} triangles (R=1, r=0.3, amp 0.05), "
                               f"{n} queries uniform in the 10%-padded AABB, flip=false, plain traversal",
                   "triangles": int(F.shape[0]), "queries": int(n)},
        "builds_per_s": 1000.0 / build_ms,
        "build_ms": build_ms,
        "scene_create_s": t_scene,
        **extra,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
