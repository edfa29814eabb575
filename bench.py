#!/usr/bin/env python3
"""Benchmark: closest-silhouette queries/s on a 2^20-triangle cut torus with
2^22 queries (plain and learned-bound), plus LBVH builds/s, on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

One step = one closest-silhouette launch over all 2^22 queries with the tree,
the queries and the outputs resident in HBM.  The learned-bound step adds the
MLP inference that produces each query's seed radius.  Prints ONE JSON line.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per closest-silhouette query of this traversal, from the
# kernel's own visit counters (one extra launch with counters on):
#   in 12 B (point) + out 8 B (distance + silhouette id)
#   + nodes visited * 64 B (node record: both child boxes and references)
#   + cone records read * 48 B (both child cones and radii)
#   + leaves evaluated * 32 B (edge range words of the leaf record)
#   + silhouette-edge tests * 32 B (endpoints, id, flags; the 32 B of face
#     normals read only for edges within the current bound are not counted)
S_NODE, S_CONE, S_LEAF, S_EDGE, S_IN, S_OUT = 64, 48, 32, 32, 12, 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--queries", type=int, default=1 << 22)
    ap.add_argument("--nu", type=int, default=1024)
    ap.add_argument("--nv", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1 << 15)
    ap.add_argument("--build-reps", type=int, default=5)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-build-scaling", action="store_true",
                    help="skip the builds of the 64-mesh batch and of an 8.4M-triangle torus")
    ap.add_argument("--input-order", action="store_true", help="also time the kernel over unordered queries")
    ap.add_argument("--sweep", action="store_true", help="also time block size x stack depth launch shapes")
    ap.add_argument("--workload", default="silhouette", choices=["silhouette", "small", "batched", "wost", "mlp"],
                    help="silhouette: the headline (1M-triangle torus, 2^22 queries); small: 20,480-triangle "
                         "icosphere, 65,536 closest-point queries; batched: 64 uneven meshes in one launch; "
                         "wost: walk on stars, 2^20 walks; mlp: learned-bound MLP training and inference")
    ap.add_argument("--pmc-mode", action="store_true",
                    help="short run for hardware-counter passes: 2 silhouette launches + a 256 MiB copy")
    ap.add_argument("--train-points", type=int, default=1 << 22)
    ap.add_argument("--train-epochs", type=int, default=20)
    ap.add_argument("--tau", type=float, default=0.95, help="expectile of the learned bound (0.5: squared error)")
    ap.add_argument("--seed-rel", type=float, default=0.25)
    ap.add_argument("--seed-abs", type=float, default=0.0)
    return ap.parse_args()


def host_cpu():
    """This host's CPU model, logical CPU count and the CPUs this process may
    run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "cpus_allowed": allowed}


def omp_threads():
    """Threads for the CPU baseline: OMP_NUM_THREADS when set (the GPU host
    sets it to this process's CPU share), else every CPU this process may use."""
    return int(os.environ.get("OMP_NUM_THREADS", "0")) or host_cpu()["cpus_allowed"]


def _median_rate(fn, n, reps):
    fn(min(n, 256))  # warm-up
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn(n)
        times.append(time.perf_counter() - t0)
    return n / float(np.median(times)), times


def cpu_baseline(V, F, q, sample, sample_1t=4096, reps=5):
    """The CPU oracle (plain-C restatement of the reference algorithm, OpenMP)
    on bounded samples of the same queries, on this host's cores: the median
    of `reps` runs after a warm-up, with all allowed threads and with one.
    Also returns the oracle's answers for the sample, for the agreement check."""
    import oracle

    threads = omp_threads()
    t0 = time.perf_counter()
    tree = oracle.build_tree(V, F)
    t_build = time.perf_counter() - t0
    qs = q[:sample]
    rate, times = _median_rate(lambda m: oracle.closest_silhouette(tree, qs[:m], nthreads=threads), sample, reps)
    rate1, _ = _median_rate(lambda m: oracle.closest_silhouette(tree, qs[:m], nthreads=1), sample_1t, 3)
    oe, od = oracle.closest_silhouette(tree, qs, nthreads=threads)
    info = host_cpu()
    return {
        "value": rate,
        "unit": "queries/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample} of the {len(q)} closest-silhouette queries (same seeds), median of {reps} runs after a "
                  f"warm-up; oracle/snch_oracle.c with OpenMP on {threads} threads",
        "one_thread_value": rate1,
        "one_thread_sample": f"{sample_1t} queries, median of 3 runs",
        "runs_s": times,
        "build_s": t_build,
        **info,
    }, tree, oe, od


def agreement(tree, V, q_sample, oe, od, ge, gd):
    """GPU answers against the oracle's on the same queries: distances bit for
    bit; silhouette ids equal, or different ids at exactly the same FP32
    distance (the library keeps the smallest id, the reference the last
    visited)."""
    import oracle

    ge = ge.view(np.uint32)
    dist_equal = bool(np.array_equal(gd, od))
    diff = ge != oe
    ties_ok = True
    if diff.any():
        a = oracle.edge_sil_dist(V, tree.slots, ge[diff], q_sample[diff])
        b = oracle.edge_sil_dist(V, tree.slots, oe[diff], q_sample[diff])
        ties_ok = bool(np.array_equal(a, b))
    return {"queries": int(len(od)), "distances_bit_exact": dist_equal, "ids_equal": int((~diff).sum()),
            "ids_tied": int(diff.sum()), "all_agree": dist_equal and ties_ok}


PMC_FILE = os.path.join(ROOT, "profiles", "pmc_silhouette_torus1M.md")
COUNTS_FILE = os.path.join(ROOT, "tests", "golden", "traversal_counts.json")


def reference_counts():
    """Per-query visit counts of the reference traversal order on this
    workload (oracle/traversal_counts.py), or None."""
    if not os.path.exists(COUNTS_FILE):
        return None
    with open(COUNTS_FILE) as f:
        return json.load(f)


def traffic_from_profile():
    """HBM bytes per launch of the silhouette kernel from the committed
    counter summary (FETCH_SIZE and WRITE_SIZE passes, each scaled by what the
    counter reads for a copy of known size); None when absent."""
    if not os.path.exists(PMC_FILE):
        return None
    with open(PMC_FILE) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def pmc_mode(sc, q, n, S, torch):
    """A few launches for `rocprofv3 --pmc`: the ordered silhouette kernel
    twice, then a device copy of 256 MiB (reads and writes 2^28 bytes each)
    that calibrates the byte counters."""
    order = sc.query_order(q)
    out_e = torch.empty(n, dtype=torch.int32, device=q.device)
    out_d = torch.empty(n, dtype=torch.float32, device=q.device)
    kopts = S.QueryOpts(0, 0, None, order.data_ptr())
    for _ in range(2):
        S.check(S.lib().snch_closest_silhouette_ex(sc._h, S._ptr(q), n, 0, None, S._ptr(out_e), S._ptr(out_d),
                                                   ctypes.byref(kopts), S._stream()))
    src = torch.ones(1 << 26, dtype=torch.float32, device=q.device)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"pmc_mode": True, "silhouette_launches": 2, "copy_bytes": src.numel() * 4}))


def run_timed(torch, fn, steps, warmup, stream=None):
    """Wall time and HIP-event time per call of `steps` calls after `warmup`
    calls; events on the stream the kernels are launched on."""
    stream = stream or torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        fn()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall / steps, ev0.elapsed_time(ev1) / steps


def base_line(metric, value, unit, args, ms_per_step, config, data, dtype="f32", higher=True):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms_per_step, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": data, "config": config}


def workload_small(args):
    """Plain method, small: 20,480-triangle noised icosphere, LBVH build, then
    65,536 closest-point queries uniform in [-1.5, 1.5]^3."""
    import torch

    import oracle
    import snch_lbvh_amd as S
    from snch_lbvh_amd import workloads as W

    V, F, q_np = W.config_small(1 << 16)
    sc = S.Scene(V, F)
    ms = []
    for _ in range(args.build_reps + 1):
        sc.build_bvh()
        ms.append(sc.build_ms())
    build_ms = float(np.median(ms[1:]))
    q = torch.from_numpy(q_np).cuda()
    n = q.shape[0]
    idx = torch.empty(n, dtype=torch.int32, device=q.device)
    dist = torch.empty(n, dtype=torch.float32, device=q.device)
    opts = S.QueryOpts(0, 0, None, None)

    def step():
        S.check(S.lib().snch_closest_point_ex(sc._h, S._ptr(q), n, S._ptr(idx), S._ptr(dist), ctypes.byref(opts),
                                              S._stream()))

    wall, ev = run_timed(torch, step, args.steps, args.warmup)
    # end to end in the reference's protocol: allocate, copy in, query, copy out
    def e2e():
        qq = torch.from_numpy(q_np).cuda()
        i2, d2 = sc.closest_point(qq)
        return i2.cpu(), d2.cpu()

    e2e_wall, _ = run_timed(torch, e2e, 8, 1)
    tree = oracle.build_tree(V, F)
    th = omp_threads()
    t0 = time.perf_counter()
    oi, od = oracle.closest_point(tree, q_np, nthreads=th)
    cpu_s = time.perf_counter() - t0
    agree = float((dist.cpu().numpy() == od).mean())
    line = base_line("closest-point queries/sec (small: 20,480-triangle icosphere)", n / wall, "queries/s", args,
                     wall * 1e3, {"workload": "icosphere level 5 (20,480 triangles, N(0,0.02) radial noise), 65,536 "
                                              "closest-point queries uniform in [-1.5,1.5]^3", "triangles": len(F),
                                  "queries": n},
                     "synthetic (seeded icosphere, uniform queries)")
    line.update({"kernel_ms": ev, "end_to_end_qps": n / e2e_wall, "build_ms": build_ms, "builds_per_s": 1e3 / build_ms,
                 "distance_agreement_with_oracle": agree,
                 "cpu_baseline": {"value": n / cpu_s, "unit": "queries/s", "cores": th, "kind": "port",
                                  "sample": f"all {n} queries, oracle/snch_oracle.c with OpenMP"}})
    print(json.dumps(line))


def workload_batched(args):
    """64 noised spheres / cut tori, triangles log-uniform in [1e3, 5e5],
    queries log-uniform in [2^12, 2^20] per mesh: one build and one
    persistent closest-silhouette launch for all, autotuned; against one
    scene and one launch per mesh."""
    import torch

    import snch_lbvh_amd as S
    from snch_lbvh_amd import workloads as W

    meshes, q_np, q_off = W.batched_meshes()
    t0 = time.perf_counter()
    b = S.BatchScene(meshes)
    t_create = time.perf_counter() - t0
    ms = []
    for _ in range(args.build_reps + 1):
        b.build_bvh()
        ms.append(b.build_ms())
    build_ms = float(np.median(ms[1:]))
    q = torch.from_numpy(q_np).cuda()
    n = q.shape[0]
    sil = torch.empty(n, dtype=torch.int32, device=q.device)
    dist = torch.empty(n, dtype=torch.float32, device=q.device)
    qo = np.ascontiguousarray(q_off, np.uint32)

    def launch(tune):
        opts = b._opts(tune, 0)
        S.check(S.lib().snch_batch_closest_silhouette(b._h, S._ptr(q), qo.ctypes.data, 0, None, S._ptr(sil),
                                                      S._ptr(dist), ctypes.byref(opts), S._stream()))

    wall0, ev0 = run_timed(torch, lambda: launch(None), max(2, args.steps // 3), 1)
    t0 = time.perf_counter()
    tune = b.autotune(q, q_off, kind="silhouette", sample_per_mesh=0)
    t_tune = time.perf_counter() - t0
    wall, ev = run_timed(torch, lambda: launch(tune), args.steps, args.warmup)
    ref = (sil.clone(), dist.clone())
    # one scene and one launch per mesh, the way the reference is used
    singles = [S.Scene(V, F).build_bvh() for V, F in meshes]
    outs = []

    def per_mesh():
        outs.clear()
        for m, sc in enumerate(singles):
            a, z = int(q_off[m]), int(q_off[m + 1])
            outs.append(sc.closest_silhouette(q[a:z]))

    wall_pm, _ = run_timed(torch, per_mesh, max(2, args.steps // 3), 1)
    same = all(torch.equal(o[1], ref[1][int(q_off[m]):int(q_off[m + 1])]) for m, o in enumerate(outs))
    line = base_line("batched closest-silhouette queries/sec (64 uneven meshes, one persistent launch)", n / wall,
                     "queries/s", args, wall * 1e3,
                     {"workload": f"64 meshes (spheres / cut tori), {sum(len(f) for _, f in meshes)} triangles, "
                                  f"{n} queries; autotuned", "triangles": int(sum(len(f) for _, f in meshes)),
                      "queries": int(n)}, "synthetic (seeded meshes and queries)")
    line.update({"kernel_ms": ev, "default_config_qps": n / wall0, "tuned": tune, "autotune_s": t_tune,
                 "per_mesh_launches_qps": n / wall_pm, "per_mesh_results_identical": bool(same),
                 "batch_build_ms": build_ms, "batch_builds_per_s": 1e3 / build_ms, "scene_create_s": t_create})
    print(json.dumps(line))


def workload_wost(args):
    """Walk on stars on the 1M-triangle torus: Dirichlet (g = x y) where the
    triangle centroid has z >= 0, zero Neumann flux elsewhere; 2^20 walks from
    uniform interior points, eps 1e-3, at most 64 steps, step loop replayed
    as a captured hipGraph."""
    import torch

    import snch_lbvh_amd as S
    from snch_lbvh_amd import workloads as W

    V, F = W.cut_torus(args.nu, args.nv, seed=3)
    FD, FN = W.split_dirichlet_neumann(V, F)
    sd = S.Scene(V, FD).build_bvh()
    sn = S.Scene(V, FN).build_bvh()
    n = 1 << 20
    x0 = torch.from_numpy(W.torus_interior_points(n, seed=8)).cuda()
    ws = S.WoSt(sd, sn, max_walks=n)
    out = torch.empty(n, dtype=torch.float32, device=x0.device)
    res = {}

    def solve(graph):
        res["r"] = ws.solve(x0, max_steps=64, eps=1e-3, seed=11, use_graph=graph, out=out)

    wall_e, _ = run_timed(torch, lambda: solve(False), 1, 1)
    eager = out.clone()
    wall, ev = run_timed(torch, lambda: solve(True), max(2, args.steps // 3), 1)
    same = bool(torch.equal(out, eager))
    line = base_line("walk-on-stars walks/sec (1M-triangle torus, 2^20 walks, <= 64 steps)", n / wall, "walks/s",
                     args, wall * 1e3,
                     {"workload": "WoSt Laplace, cut torus 1,048,576 triangles split Dirichlet (z >= 0) / Neumann, "
                                  "2^20 walks, eps 1e-3, max 64 steps, hipGraph", "walks": n,
                      "dirichlet_triangles": len(FD), "neumann_triangles": len(FN)},
                     "synthetic (seeded torus and starts)")
    line.update({"kernel_ms": ev, "eager_walks_per_s": n / wall_e, "graph_equals_eager": same,
                 "live_after": int(res["r"][1]), "mean_estimate": float(out.mean())})
    print(json.dumps(line))


def workload_mlp(args):
    """Learned distance bound: 2^24 training points (half uniform in the
    padded AABB, half near-surface with N(0, 0.05) offsets) labelled by the
    exact GPU closest point on the 1M-triangle torus; training epochs and
    2^22-point inference on BF16 MFMA; error against the FP64 brute force."""
    import torch

    import oracle
    import snch_lbvh_amd as S
    from snch_lbvh_amd import workloads as W

    V, F = W.cut_torus(args.nu, args.nv, seed=3)
    lo, hi = W.padded_box(V)
    sc = S.Scene(V, F).build_bvh()
    pts = torch.from_numpy(W.mlp_training_points(V, F, 1 << 24, seed=9)).cuda()
    _, lab = sc.closest_point(pts)
    mlp = S.LearnedBound(lo.astype(np.float32), hi.astype(np.float32), sigma=4.0, seed=1)
    epochs = 4
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = mlp.train(pts, lab, batch=4096, epochs=epochs, lr=2e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    qi = torch.from_numpy(W.uniform_points(1 << 22, lo, hi, seed=4)).cuda()
    pred = torch.empty(qi.shape[0], dtype=torch.float32, device=qi.device)
    wall, ev = run_timed(torch, lambda: mlp.predict(qi, out=pred), args.steps, args.warmup)
    flop_fwd = 2.0 * (64 * 128 + 3 * 128 * 128 + 128)  # per point
    n_inf = qi.shape[0]
    infer_tflops = n_inf * flop_fwd / (ev * 1e-3) / 1e12
    train_tflops = epochs * (1 << 24) * 3 * flop_fwd / t_train / 1e12
    # error against the FP64 brute force on a sample
    ns = 512
    qs = W.uniform_points(ns, lo, hi, seed=77)
    _, d64 = oracle.bf_closest_point(V, F, qs, nthreads=omp_threads())
    p = mlp.predict(qs).cpu().numpy()
    rel = np.abs(p - d64) / np.maximum(d64, 1e-6)
    line = base_line("learned-bound MLP inference points/sec (BF16 MFMA, 2^22 points)", n_inf / wall, "points/s",
                     args, wall * 1e3,
                     {"workload": "3->64 Fourier (sigma 4) -> 4x128 SiLU -> 1 MLP; train 2^24 points labelled by "
                                  "exact closest point on the 1M-triangle torus; infer 2^22", "train_points": 1 << 24,
                      "infer_points": n_inf}, "synthetic (seeded torus, sampled points)", dtype="bf16")
    line.update({"kernel_ms": ev, "train_s": t_train, "train_epochs": epochs,
                 "train_points_per_s": epochs * (1 << 24) / t_train,
                 "loss_first_last": [float(losses[0]), float(losses[-1])],
                 "roofline": {"bound": "mfma", "achieved": infer_tflops, "peak": 2500.0, "unit": "TFLOP/s",
                              "frac": infer_tflops / 2500.0, "traffic": None, "kernel": "k_mlp_infer",
                              "train_tflops": train_tflops},
                 "error_vs_fp64": {"sample": ns, "median_rel": float(np.median(rel)),
                                   "p90_rel": float(np.quantile(rel, 0.9)), "max_rel": float(rel.max())}})
    print(json.dumps(line))


# Algorithmic bytes of each build phase per primitive (reference payload widths,
# no padding): leaf generation 188 + scene box 24 + Morton 32; the stable sort
# of 30-bit codes in four 8-bit passes, 4 + 16 per pass; Karras 28; the gather
# of leaf boxes and cones into sorted order 88 plus the AABB + cone refit 164.
BUILD_PHASE_BYTES = {"leaves": 244, "sort": 68, "karras": 28, "refit": 252}


def build_phase_roofline(sc, n_prims, reps):
    """Median GPU time of each build phase over `reps` builds timed with
    events between the phases, against the HBM peak."""
    ph = []
    for _ in range(max(reps, 1) + 1):
        sc.build_bvh(time_phases=True)
        ph.append(sc.build_phases())
    out = {}
    for k, per in BUILD_PHASE_BYTES.items():
        ms = float(np.median([p[k + "_ms"] for p in ph[1:]]))
        gbs = per * n_prims / (ms * 1e-3) / 1e9 if ms > 0 else None
        out[k] = {"ms": ms, "bytes_per_prim": per, "achieved_GBps": gbs,
                  "frac": gbs / HBM_PEAK_GBS if gbs else None}
    sc.build_bvh()
    return out


def build_scaling(reps):
    """Builds/s and the phase roofline where launch latency weighs less: the
    64-mesh batch (4.76M triangles, one build) and an 8.4M-triangle torus."""
    import snch_lbvh_amd as S
    from snch_lbvh_amd import workloads as W

    out = {}
    meshes, _, _ = W.batched_meshes()
    V8, F8 = W.cut_torus(2896, 1448, seed=3)
    for name, sc, n in (("batch_64_meshes", S.BatchScene(meshes), int(sum(len(f) for _, f in meshes))),
                        ("torus_8M", S.Scene(V8, F8), int(F8.shape[0]))):
        ms = []
        for _ in range(reps + 1):
            sc.build_bvh()
            ms.append(sc.build_ms())
        build_ms = float(np.median(ms[1:]))
        out[name] = {"triangles": n, "build_ms": build_ms, "builds_per_s": 1e3 / build_ms,
                     "whole_build_frac": sum(BUILD_PHASE_BYTES.values()) * n / (build_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "phases": build_phase_roofline(sc, n, reps)}
        del sc
    return out


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: the library has no multi-GPU path "
              "(query batches would shard across GPUs without communication)", file=sys.stderr)
        sys.exit(2)
    if args.workload != "silhouette":
        {"small": workload_small, "batched": workload_batched, "wost": workload_wost,
         "mlp": workload_mlp}[args.workload](args)
        return
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
    build_roofline = build_phase_roofline(sc, F.shape[0], args.build_reps)
    build_larger = None if args.no_build_scaling else build_scaling(args.build_reps)

    q = torch.from_numpy(q_np).to(dev)
    n = q.shape[0]
    if args.pmc_mode:
        pmc_mode(sc, q, n, S, torch)
        return
    out_e = torch.empty(n, dtype=torch.int32, device=dev)
    out_d = torch.empty(n, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    def step():
        S.check(S.lib().snch_closest_silhouette(sc._h, S._ptr(q), n, 0, None, S._ptr(out_e), S._ptr(out_d),
                                                S._stream()))

    def run_timed(fn, steps, warmup):
        """Wall time and event time of `steps` calls, after `warmup` calls;
        the events are recorded on the stream the kernels are launched on."""
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(steps):
            fn()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall, ev0.elapsed_time(ev1) / steps

    wall, step_ms = run_timed(step, args.steps, args.warmup)
    ms_per_step = wall * 1000.0 / args.steps
    value = n * args.steps / wall

    # the step = Morton ordering of the queries + the traversal kernel; time
    # each alone (the kernel with the order supplied by the caller)
    order = sc.query_order(q)
    kopts = S.QueryOpts(0, 0, None, order.data_ptr())

    def kstep():
        S.check(S.lib().snch_closest_silhouette_ex(sc._h, S._ptr(q), n, 0, None, S._ptr(out_e), S._ptr(out_d),
                                                   ctypes.byref(kopts), S._stream()))

    _, kernel_ms = run_timed(kstep, args.steps, 1)
    # launch shapes of the same kernel (results identical): block size x LDS stack depth
    sweep = {}
    if args.sweep:
        for bs in (64, 128, 256):
            for depth in (8, 16, 24):
                o2 = S.QueryOpts(bs, 0, None, order.data_ptr(), depth, 0)
                _, ms2 = run_timed(lambda: S.check(S.lib().snch_closest_silhouette_ex(
                    sc._h, S._ptr(q), n, 0, None, S._ptr(out_e), S._ptr(out_d), ctypes.byref(o2), S._stream())), 2, 1)
                sweep[f"{bs}x{depth}"] = ms2
        cp_i = torch.empty_like(out_e)
        cp_d = torch.empty_like(out_d)
        for bs in (64, 256):
            for depth in (8, 16):
                o3 = S.QueryOpts(bs, 0, None, order.data_ptr(), depth, 0)
                _, ms3 = run_timed(lambda: S.check(S.lib().snch_closest_point_ex(
                    sc._h, S._ptr(q), n, S._ptr(cp_i), S._ptr(cp_d), ctypes.byref(o3), S._stream())), 2, 1)
                sweep[f"closest_point_{bs}x{depth}"] = ms3
    perm_buf = torch.empty(n, dtype=torch.int32, device=dev)
    _, order_ms = run_timed(lambda: S.check(S.lib().snch_query_order(sc._h, S._ptr(q), n, S._ptr(perm_buf),
                                                                     S._stream())), args.steps, 1)
    input_order_ms = None
    if args.input_order:  # the same kernel over the queries in input order, for comparison
        iopts = S.QueryOpts(0, 2, None, None)
        _, input_order_ms = run_timed(lambda: S.check(S.lib().snch_closest_silhouette_ex(
            sc._h, S._ptr(q), n, 0, None, S._ptr(out_e), S._ptr(out_d), ctypes.byref(iopts), S._stream())), 2, 1)

    # visit counters of the same launch (one extra launch, counters on)
    stats = torch.zeros(5, dtype=torch.int64, device=dev)
    e_chk, d_chk = sc.closest_silhouette(q, stats=stats, order=order)
    assert torch.equal(e_chk, out_e) and torch.equal(d_chk, out_d)
    box, leaf, edge, node, cone = (v / n for v in stats.tolist())
    per_q = S_IN + S_OUT + node * S_NODE + cone * S_CONE + leaf * S_LEAF + edge * S_EDGE
    achieved = per_q * n / (kernel_ms * 1e-3) / 1e9
    traffic = traffic_from_profile()
    # the same rate with the reference traversal's visit counts (measured once
    # by the oracle on this workload) at the reference's payload widths: box
    # tests x (24 + 8 + 20) B, leaves x 40 B, edges x (32 + 48) B
    ref_counts = reference_counts()
    ref_bytes = None
    if ref_counts:
        c = ref_counts["silhouette"]
        ref_bytes = 13 + 8 + c["box_tests"] * (24 + 8 + 20) + c["leaf_evals"] * 40 + c["edge_tests"] * (32 + 48)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "bytes_per_query": per_q,
                "definition": "algorithmic bytes = this kernel's own per-query visits (counters of the same launch) "
                              "x record sizes as read; the reference traversal's visits would give "
                              "bytes_per_query_reference_order (it tests ~4x more boxes)",
                "bytes_per_query_reference_order": ref_bytes,
                "traffic_frac": (traffic / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
                "kernel": "k_sil_persist<16, false, 3, false>", "kernel_ms": kernel_ms, "step_ms_events": step_ms,
                "query_order_ms": order_ms, "input_order_kernel_ms": input_order_ms, "shape_sweep_ms": sweep,
                "visits_per_query": {"nodes": node, "cone_records": cone, "leaves": leaf, "edge_tests": edge}}

    # learned-bound form: seed radius from the Fourier MLP (BF16 MFMA), then
    # the seeded exact traversal; a step is inference + query, results equal
    # the plain step's bit for bit
    learned = None
    if not args.no_learned:
        lo32, hi32 = lo.astype(np.float32), hi.astype(np.float32)
        tp = torch.from_numpy(W.uniform_points(args.train_points, lo, hi, seed=21)).to(dev)
        _, td = sc.closest_silhouette(tp)
        mlp = S.LearnedBound(lo32, hi32, sigma=4.0, seed=1).set_loss(args.tau)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        losses = mlp.train(tp, td, batch=4096, epochs=args.train_epochs, lr=3e-3)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        seed_r = torch.empty(n, dtype=torch.float32, device=dev)
        out_e2 = torch.empty_like(out_e)
        out_d2 = torch.empty_like(out_d)

        def seeded(seed_tensor, in_place=0):
            # unsettled queries re-queued for whole waves (the default), or searched again in place
            opts = S.QueryOpts(0, 0, None, None, 0, 0, 0, in_place)
            S.check(S.lib().snch_closest_silhouette_ex(sc._h, S._ptr(q), n, 0, S._ptr(seed_tensor), S._ptr(out_e2),
                                                       S._ptr(out_d2), ctypes.byref(opts), S._stream()))

        # margin sweep: seed = prediction * (1 + rel) + abs * scene diagonal
        margins = {}
        for rel, ab in ((0.05, args.seed_abs), (0.1, args.seed_abs), (0.25, args.seed_abs), (0.05, 0.002), (0.1, 0.002)):
            def lstep(rel=rel, ab=ab):
                mlp.seed_radius(q, rel=rel, abs_margin=ab, out=seed_r)
                seeded(seed_r)

            lwall, _ = run_timed(lstep, max(2, args.steps // 3), 1)
            assert torch.equal(out_e2, out_e) and torch.equal(out_d2, out_d), "seeded result differs"
            fb = float(((out_d > seed_r) | ~torch.isfinite(out_d)).float().mean())
            margins[f"{rel}/{ab}"] = {"rel": rel, "abs": ab, "ms_per_step": lwall * 1000.0 / max(2, args.steps // 3),
                                      "fallback_fraction": fb}
        best = min(margins, key=lambda r: margins[r]["ms_per_step"])
        best_rel, best_abs = margins[best]["rel"], margins[best]["abs"]

        def lstep(in_place=0):
            mlp.seed_radius(q, rel=best_rel, abs_margin=best_abs, out=seed_r)
            seeded(seed_r, in_place)

        lwall, lms = run_timed(lstep, args.steps, args.warmup)
        assert torch.equal(out_e2, out_e) and torch.equal(out_d2, out_d), "seeded result differs"
        pwall, _ = run_timed(lambda: lstep(1), max(2, args.steps // 3), 1)
        assert torch.equal(out_e2, out_e) and torch.equal(out_d2, out_d), "seeded result differs"
        st2 = torch.zeros(5, dtype=torch.int64, device=dev)
        sc.closest_silhouette(q, seed_radius=seed_r, stats=st2)
        fallback = float(((out_d > seed_r) | ~torch.isfinite(out_d)).float().mean())
        # the bound a perfect predictor would give: the exact answer itself
        ideal = torch.where(torch.isfinite(out_d), out_d, torch.full_like(out_d, -1.0))
        iwall, _ = run_timed(lambda: seeded(ideal), max(2, args.steps // 3), 1)
        st3 = torch.zeros(5, dtype=torch.int64, device=dev)
        sc.closest_silhouette(q, seed_radius=ideal, stats=st3)
        learned = {"value": n * args.steps / lwall, "ms_per_step": lwall * 1000.0 / args.steps,
                   "train_s": t_train, "train_points": args.train_points, "train_epochs": args.train_epochs,
                   "loss_first_last": [float(losses[0]), float(losses[-1])],
                   "nodes_per_query": st2[3].item() / n, "seed_rel": best_rel, "seed_abs": best_abs,
                   "fallback_fraction": fallback, "margin_sweep": margins, "tau": args.tau,
                   "unsettled": "re-queued for whole waves of the same launch",
                   "in_place_ms_per_step": pwall * 1000.0 / max(2, args.steps // 3),
                   "exact_seed": {"value": n * max(2, args.steps // 3) / iwall, "nodes_per_query": st3[3].item() / n}}

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

    # end to end in the reference's protocol (benchmark.cu:393-416): device
    # allocation, host-to-device copy of the queries, the query, copy of the
    # results back; mean of 32 after one warm-up.  Not `value`, which has the
    # inputs resident in HBM.
    def e2e():
        qq = torch.from_numpy(q_np).to(dev)
        e2, d2 = sc.closest_silhouette(qq)
        return e2.cpu(), d2.cpu()

    e2e_wall, _ = run_timed(e2e, 32, 1)
    extra["end_to_end_qps"] = n * 32 / e2e_wall
    extra["end_to_end_protocol"] = "alloc + H2D of the queries + query + D2H of both outputs, mean of 32 after 1 warm-up"

    cpu = None
    agree = None
    if not args.no_cpu_baseline:
        cpu, otree, oe, od = cpu_baseline(V, F, q_np, args.cpu_sample)
        m = len(od)
        agree = agreement(otree, V, q_np[:m], oe, od, out_e[:m].cpu().numpy(), out_d[:m].cpu().numpy())
        if not agree["all_agree"]:
            print(json.dumps({"error": "GPU results differ from the oracle", "agreement": agree}), file=sys.stderr)
            sys.exit(1)

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
        "build_roofline": build_roofline,
        "build_larger": build_larger,
        "scene_create_s": t_scene,
        "learned_bound": learned,
        **extra,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "agreement_with_oracle": agree,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
