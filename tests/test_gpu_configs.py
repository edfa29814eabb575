"""The benchmark configurations at their full sizes, on the GPU, checked
against the oracle on seeded subsamples and through size-independent
properties on the full sets:

  C2  2^20-triangle cut torus, 2^22 queries: closest point, closest
      silhouette (both orientations), ray (closest and any hit) and
      SampleTriangleInSphere;
  C3  64 uneven meshes (4.8M triangles, 12M queries) in one batch;
  C4  walk on stars, 2^20 walks on the 2^20-triangle torus;
  C5  the learned bound: training on 2^24 points, seeded search over 2^22.

Bars: trees bit-exact; distances, hit t and samples bit-exact against the
oracle; indices equal except exact distance ties (the library keeps the
smallest index, the reference the last or first visited); FP64 brute force
agreement of the returned primitive's distance within 1e-6 relative."""
import numpy as np
import pytest

from snch_lbvh_amd import workloads as W

pytestmark = pytest.mark.gpu

SUB = 16384  # oracle subsample per query type


@pytest.fixture(scope="module")
def S():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import snch_lbvh_amd as S

    S.lib()
    return S


def _ties_ok(gi, oi, gd, od, dist_of):
    """Distances bit-exact; where the indices differ, both candidates are at
    exactly the same FP32 distance."""
    np.testing.assert_array_equal(gd, od)
    diff = gi != oi
    if diff.any():
        np.testing.assert_array_equal(dist_of(gi[diff], diff), dist_of(oi[diff], diff))
    return int(diff.sum())


@pytest.fixture(scope="module")
def c2(S, oracle):
    import torch

    V, F = W.cut_torus(1024, 512, seed=3)
    lo, hi = W.padded_box(V)
    q = W.uniform_points(1 << 22, lo, hi, seed=4)
    sc = S.Scene(V, F).build_bvh()
    tree = oracle.build_tree(V, F)
    sub = np.sort(np.random.default_rng(99).choice(len(q), SUB, replace=False))
    return {"V": V, "F": F, "q": q, "qd": torch.from_numpy(q).cuda(), "sc": sc, "tree": tree, "sub": sub}


def test_c2_tree_bit_exact(c2):
    g = c2["sc"].export()
    t = c2["tree"]
    for k, ov in (("sorted_codes", t.sorted_codes), ("sorted_idx", t.sorted_idx), ("nodes", t.nodes),
                  ("boxes", t.boxes), ("cones", t.cones)):
        np.testing.assert_array_equal(g[k], ov, err_msg=k)


def test_c2_closest_point(c2, oracle):
    V, F, q, sub = c2["V"], c2["F"], c2["q"], c2["sub"]
    sc = c2["sc"]
    gi, gd = (t.cpu().numpy() for t in sc.closest_point(c2["qd"]))
    gi = gi.view(np.uint32)
    oi, od = oracle.closest_point(c2["tree"], q[sub])
    qs = q[sub]
    _ties_ok(gi[sub], oi, gd[sub], od, lambda tri, m: oracle.tri_dist(V, F, tri, qs[m]))
    # input order gives the Morton-order results bit for bit over all 2^22 queries
    ii, idd = (t.cpu().numpy() for t in sc.closest_point(c2["qd"], coherent=2))
    np.testing.assert_array_equal(ii.view(np.uint32), gi)
    np.testing.assert_array_equal(idd, gd)
    # FP64 brute force on 256 queries: the returned triangle is the nearest
    # within 1e-6 relative
    b = sub[:256]
    bi, bd = oracle.bf_closest_point(V, F, q[b])
    d64 = oracle.tri_dist(V, F, gi[b], q[b], f64=True)
    assert np.all(d64 - bd <= 1e-6 * np.maximum(bd, 1e-12))
    assert np.all(np.abs(gd[b] - bd) <= 1e-6 * bd + 1e-7)


@pytest.mark.parametrize("flip", [False, True])
def test_c2_closest_silhouette(c2, oracle, flip):
    V, q, sub = c2["V"], c2["q"], c2["sub"]
    sc, tree = c2["sc"], c2["tree"]
    ge, gd = (t.cpu().numpy() for t in sc.closest_silhouette(c2["qd"], flip=flip))
    ge = ge.view(np.uint32)
    oe, od = oracle.closest_silhouette(tree, q[sub], flip=flip)
    qs = q[sub]
    _ties_ok(ge[sub], oe, gd[sub], od, lambda e, m: oracle.edge_sil_dist(V, tree.slots, e, qs[m], flip=flip))
    ie, idd = (t.cpu().numpy() for t in sc.closest_silhouette(c2["qd"], flip=flip, coherent=2))
    np.testing.assert_array_equal(ie.view(np.uint32), ge)
    np.testing.assert_array_equal(idd, gd)
    # seeded search (exact, loose, too small, none) equals the plain search on all 2^22
    import torch

    d = torch.from_numpy(gd).cuda()
    seed = torch.where(torch.isfinite(d), d * 1.05, torch.full_like(d, 0.05))
    seed[::5] = 0.0
    se, sdd = (t.cpu().numpy() for t in sc.closest_silhouette(c2["qd"], flip=flip, seed_radius=seed))
    np.testing.assert_array_equal(se.view(np.uint32), ge)
    np.testing.assert_array_equal(sdd, gd)
    # FP64 brute force on 256 queries: the returned edge's distance is the minimum
    b = sub[:256]
    be, bd = oracle.bf_closest_silhouette(V, tree.slots, q[b], flip=flip)
    fin = np.isfinite(bd)
    assert np.array_equal(np.isfinite(gd[b]), fin)
    d64 = oracle.edge_sil_dist64(V, tree.slots, ge[b][fin], q[b][fin], flip=flip)
    assert np.all(d64 - bd[fin] <= 1e-6 * np.maximum(bd[fin], 1e-12))


def test_c2_ray(c2, oracle):
    V, F, q, sub = c2["V"], c2["F"], c2["q"], c2["sub"]
    sc = c2["sc"]
    d = W.unit_directions(len(q), 5)
    hit, t, uv, idx = (x.cpu().numpy() for x in sc.ray_intersect(c2["qd"], d))
    oh, ot, ouv, oidx = oracle.ray(c2["tree"], q[sub], d[sub])
    np.testing.assert_array_equal(hit[sub], oh)
    h = oh
    np.testing.assert_array_equal(t[sub][h], ot[h])
    same = idx.view(np.uint32)[sub] == oidx
    np.testing.assert_array_equal(uv[sub][h & same], ouv[h & same])
    ahit = sc.ray_intersect(c2["qd"], d, any_hit=True)[0].cpu().numpy()
    np.testing.assert_array_equal(ahit, hit)
    # FP64 brute force on 256 rays
    b = sub[:256]
    bi, bt = oracle.bf_ray(V, F, q[b], d[b])
    fin = np.isfinite(bt)
    assert np.array_equal(hit[b], fin)
    assert np.all(np.abs(t[b][fin] - bt[fin]) <= 1e-5 * np.maximum(bt[fin], 1e-3))


def test_c2_sample_in_sphere(c2, oracle):
    q, sub = c2["q"], c2["sub"]
    n = len(q)
    r = W.log_uniform(n, 1e-3, 0.5, 6)
    u = np.random.default_rng(7).random((n, 3)).astype(np.float32)
    gi, gp, gpt = (x.cpu().numpy() for x in c2["sc"].sample_in_sphere(c2["qd"], r, u))
    oi, op, opt = oracle.sample_in_sphere(c2["tree"], q[sub], r[sub], u[sub])
    np.testing.assert_array_equal(gi[sub], oi)
    np.testing.assert_array_equal(gp[sub], op)
    np.testing.assert_array_equal(gpt[sub], opt)
    ok = gi >= 0
    assert ok.mean() > 0.05
    assert np.all(gp[ok] > 0) and np.all(gp[~ok] == 0)


# ------------------------------------------------------------ C3
def test_c3_batched(S, oracle):
    """64 meshes built and queried in one batch: every mesh's tree equals the
    oracle's, batched results equal one scene and one launch per mesh, and a
    per-mesh subsample equals the oracle's queries."""
    import torch

    meshes, q, q_off = W.batched_meshes()
    b = S.BatchScene(meshes).build_bvh()
    qd = torch.from_numpy(q).cuda()
    ci, cd = (t.cpu().numpy() for t in b.closest_point(qd, q_off))
    se, sd = (t.cpu().numpy() for t in b.closest_silhouette(qd, q_off))
    ci, se = ci.view(np.uint32), se.view(np.uint32)
    # the learned-bound form of the batch (a seed radius per query) changes no result
    sdd = torch.from_numpy(sd).cuda()
    seed = torch.where(torch.isfinite(sdd), sdd * 1.1, torch.zeros_like(sdd))
    e2, d2 = (t.cpu().numpy() for t in b.closest_silhouette(qd, q_off, seed_radius=seed))
    np.testing.assert_array_equal(e2.view(np.uint32), se)
    np.testing.assert_array_equal(d2, sd)
    rng = np.random.default_rng(5)
    for m, (V, F) in enumerate(meshes):
        a, z = int(q_off[m]), int(q_off[m + 1])
        t = oracle.build_tree(V, F)
        g = b.export(m)
        for k, ov in (("sorted_idx", t.sorted_idx), ("nodes", t.nodes), ("boxes", t.boxes), ("cones", t.cones)):
            np.testing.assert_array_equal(g[k], ov, err_msg=f"mesh {m} {k}")
        sc = S.Scene(V, F).build_bvh()
        i1, d1 = (x.cpu().numpy() for x in sc.closest_point(qd[a:z]))
        e1, s1 = (x.cpu().numpy() for x in sc.closest_silhouette(qd[a:z]))
        np.testing.assert_array_equal(i1.view(np.uint32), ci[a:z])
        np.testing.assert_array_equal(d1, cd[a:z])
        np.testing.assert_array_equal(e1.view(np.uint32), se[a:z])
        np.testing.assert_array_equal(s1, sd[a:z])
        pick = a + np.sort(rng.choice(z - a, min(256, z - a), replace=False))
        qs = q[pick]
        oi, od = oracle.closest_point(t, qs)
        _ties_ok(ci[pick], oi, cd[pick], od, lambda tri, mm: oracle.tri_dist(V, F, tri, qs[mm]))
        oe, osd = oracle.closest_silhouette(t, qs)
        _ties_ok(se[pick], oe, sd[pick], osd, lambda e, mm: oracle.edge_sil_dist(V, t.slots, e, qs[mm]))
        del sc


# ------------------------------------------------------------ C4
def test_c4_wost(S, oracle):
    """2^20 walks on the 2^20-triangle torus (Dirichlet z >= 0, Neumann
    below), <= 64 steps, eps 1e-3: the captured graph equals the eager loop,
    every walk ends, estimates are values of g on the Dirichlet part, and the
    first 4,096 walks equal the CPU restatement's bit for bit."""
    V, F = W.cut_torus(1024, 512, seed=3)
    FD, FN = W.split_dirichlet_neumann(V, F)
    sd = S.Scene(V, FD).build_bvh()
    sn = S.Scene(V, FN).build_bvh()
    n = 1 << 20
    x0 = W.torus_interior_points(n, seed=8)
    ws = S.WoSt(sd, sn, max_walks=n)
    a, la = ws.solve(x0, max_steps=64, eps=1e-3, seed=11, use_graph=False)
    a = a.cpu().numpy()
    g, lg = ws.solve(x0, max_steps=64, eps=1e-3, seed=11, use_graph=True)
    assert la == lg == 0
    np.testing.assert_array_equal(g.cpu().numpy(), a)
    assert np.all(np.isfinite(a)) and np.all(np.abs(a) < 1.0)
    k = 4096
    o, _ = oracle.wost(oracle.build_tree(V, FD), oracle.build_tree(V, FN), x0[:k], 64, 1e-3, 11)
    np.testing.assert_array_equal(a[:k], o)


# ------------------------------------------------------------ C5
def test_c5_learned_bound(S, oracle):
    """Training on 2^24 points (uniform in the padded box and near the
    surface) labelled by the exact GPU closest point; labels checked against
    the oracle on a subsample; the loss falls; seeds from the trained network
    leave every one of 2^22 closest-silhouette results unchanged."""
    import torch

    V, F = W.cut_torus(1024, 512, seed=3)
    lo, hi = W.padded_box(V)
    sc = S.Scene(V, F).build_bvh()
    pts = torch.from_numpy(W.mlp_training_points(V, F, 1 << 24, seed=9)).cuda()
    li, lab = sc.closest_point(pts)
    sub = np.sort(np.random.default_rng(3).choice(1 << 24, 4096, replace=False))
    ps = pts[torch.from_numpy(sub).cuda()].cpu().numpy()
    _, od = oracle.closest_point(oracle.build_tree(V, F), ps)
    np.testing.assert_array_equal(lab.cpu().numpy()[sub], od)
    mlp = S.LearnedBound(lo.astype(np.float32), hi.astype(np.float32), sigma=4.0, seed=1)
    losses = mlp.train(pts, lab, batch=4096, epochs=2, lr=2e-3)
    assert np.all(np.isfinite(losses)) and losses[-1] < losses[0]
    q = torch.from_numpy(W.uniform_points(1 << 22, lo, hi, seed=4)).cuda()
    e0, d0 = sc.closest_silhouette(q)
    seed = mlp.seed_radius(q, rel=0.25, abs_margin=0.005)
    e1, d1 = sc.closest_silhouette(q, seed_radius=seed)
    assert torch.equal(e0, e1) and torch.equal(d0, d1)
