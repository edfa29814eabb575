"""GPU parity: the HIP kernels behind the C ABI against the CPU oracle.

Bars: build (Morton codes, sorted order, topology, AABBs, cones) bit-exact;
sampler bit-exact; closest point / silhouette / ray: distances bit-exact and
indices identical except where the oracle's FP32 distance of both candidates
is exactly tied (the library breaks ties by smallest index, the reference by
visit order)."""
import numpy as np
import pytest

from snch_lbvh_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import snch_lbvh_amd as S

    S.lib()
    return S


def _soup(seed=9, nv=300, nf=500):
    rng = np.random.default_rng(seed)
    return rng.random((nv, 3)).astype(np.float32), rng.integers(0, nv, size=(nf, 3)).astype(np.int32)


MESHES = {
    "ico5": lambda: W.icosphere(5, 0.02, seed=1),
    "ico2_clean": lambda: W.icosphere(2, 0.0),
    "torus": lambda: W.cut_torus(96, 48),
    "soup": _soup,
    "one": lambda: (np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]]), np.int32([[0, 1, 2]])),
    "two": lambda: (np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.5]]), np.int32([[0, 1, 2], [1, 3, 2]])),
    "three": lambda: (np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.5], [2, 0, 1]]),
                      np.int32([[0, 1, 2], [1, 3, 2], [1, 4, 3]])),
}


def _queries(V, n, seed):
    lo, hi = W.padded_box(V, 0.25)
    return W.uniform_points(n, lo, hi, seed)


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("reference_radius", [False, True])
def test_build_bit_exact(S, oracle, name, reference_radius):
    V, F = MESHES[name]()
    sc = S.Scene(V, F).build_bvh(reference_radius=reference_radius)
    g = sc.export()
    o = oracle.build_tree(V, F, conservative_radius=not reference_radius)
    slots, owned = sc.silhouettes()
    _, oslots, oowned = oracle.edge_table(F)
    np.testing.assert_array_equal(slots, oslots)
    np.testing.assert_array_equal(owned, oowned)
    np.testing.assert_array_equal(g["sorted_codes"], o.sorted_codes)
    np.testing.assert_array_equal(g["sorted_idx"], o.sorted_idx)
    np.testing.assert_array_equal(g["nodes"], o.nodes)
    np.testing.assert_array_equal(g["boxes"], o.boxes)
    np.testing.assert_array_equal(g["cones"], o.cones)


def test_build_bit_exact_large(S, oracle):
    """327,680 triangles: the radix sort's multi-block histogram scan and all
    three refit levels (subtrees of 1,024 leaves, mid sets of 16,384, the
    nodes above) run; the whole tree must equal the oracle's.
    Then 300,000 queries in Morton order against input order."""
    V, F = W.cut_torus(512, 320)
    sc = S.Scene(V, F).build_bvh()
    g = sc.export()
    o = oracle.build_tree(V, F)
    for k, ov in (("sorted_codes", o.sorted_codes), ("sorted_idx", o.sorted_idx), ("nodes", o.nodes),
                  ("boxes", o.boxes), ("cones", o.cones)):
        np.testing.assert_array_equal(g[k], ov, err_msg=k)
    q = _queries(V, 300000, 13)
    perm = sc.query_order(q).cpu().numpy()
    assert np.array_equal(np.sort(perm), np.arange(len(q)))
    a = [t.cpu().numpy() for t in sc.closest_silhouette(q, coherent=1)]
    b = [t.cpu().numpy() for t in sc.closest_silhouette(q, coherent=2)]
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_build_bit_exact_many_tops(S, oracle):
    """2,096,704 triangles: about 3,500 refit subtrees of up to 1,024 leaves
    and a few hundred mid sets; the tree must equal the oracle's with every
    layout of the top pass (boxes in LDS, boxes in memory, the climb with
    arrival counters that very large scenes use)."""
    V, F = W.cut_torus(1448, 724)
    o = oracle.build_tree(V, F)
    sc = S.Scene(V, F)
    for layout in ("auto", "memory", "climb"):
        g = sc.build_bvh(top_layout=layout).export()
        for k, ov in (("sorted_idx", o.sorted_idx), ("nodes", o.nodes), ("boxes", o.boxes), ("cones", o.cones)):
            np.testing.assert_array_equal(g[k], ov, err_msg=f"{layout}: {k}")


def test_build_bit_exact_8m(S, oracle):
    """8,386,816 triangles: about 14,000 refit subtrees, about 1,000 mid sets
    and the nodes above them; the tree must equal the oracle's."""
    V, F = W.cut_torus(2896, 1448)
    g = S.Scene(V, F).build_bvh().export()
    o = oracle.build_tree(V, F)
    for k, ov in (("sorted_idx", o.sorted_idx), ("nodes", o.nodes), ("boxes", o.boxes), ("cones", o.cones)):
        np.testing.assert_array_equal(g[k], ov, err_msg=k)
    del g, o


def test_build_phase_timing(S):
    """Timing the build's phases records four positive GPU times that add up
    to at most the whole build, and leaves the tree unchanged."""
    V, F = W.cut_torus(256, 128)
    sc = S.Scene(V, F).build_bvh()
    a = sc.export()
    assert all(v == 0.0 for v in sc.build_phases().values())
    sc.build_bvh(time_phases=True)
    ph = sc.build_phases()
    assert all(v > 0.0 for v in ph.values()), ph
    assert sum(ph.values()) <= sc.build_ms() * 1.05 + 0.01
    b = sc.export()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])


def test_rebuild_replays_identical(S, oracle):
    """Repeat builds of one scene replay a captured launch sequence (from the
    second build with the same options on); every rebuild, also after a
    change of leaf size and back, must give the oracle's tree, and queries
    after a replayed build must match queries after the first build."""
    V, F = W.cut_torus(192, 96)
    o = oracle.build_tree(V, F)
    sc = S.Scene(V, F)
    q = _queries(V, 20000, 5)
    ref = None
    for leaf in (1, 1, 1, 4, 4, 4, 1, 1):
        sc.build_bvh(leaf_size=leaf)
        if leaf == 1:
            g = sc.export()
            for k, ov in (("sorted_idx", o.sorted_idx), ("nodes", o.nodes), ("boxes", o.boxes), ("cones", o.cones)):
                np.testing.assert_array_equal(g[k], ov, err_msg=k)
        res = [t.cpu().numpy() for t in sc.closest_silhouette(q)]
        if ref is None:
            ref = res
        for x, y in zip(res, ref):
            np.testing.assert_array_equal(x, y)


def test_no_device_errors(S):
    """After builds, a large query batch and a sort, the device's sticky error
    words (traversal overflow, sort wait bound) are clear."""
    V, F = W.cut_torus(512, 256)
    sc = S.Scene(V, F).build_bvh()
    q = _queries(V, 200000, 3)
    sc.closest_silhouette(q)
    sc.closest_point(q)
    S.check(S.lib().snch_stream_synchronize(S._stream()))


def test_build_deterministic(S):
    V, F = W.cut_torus(128, 64)
    a = S.Scene(V, F).build_bvh().export()
    b = S.Scene(V, F).build_bvh().export()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])


def _check_ties(oracle, V, F, q, gi, oi, gd, od, f32dist):
    np.testing.assert_array_equal(gd, od)
    diff = gi != oi
    if diff.any():
        dg = f32dist(gi[diff], q[diff])
        do = f32dist(oi[diff], q[diff])
        np.testing.assert_array_equal(dg, do)
    return float(np.mean(diff))


@pytest.mark.parametrize("name", ["ico5", "torus", "soup", "one", "two", "three"])
def test_closest_point_parity(S, oracle, name):
    V, F = MESHES[name]()
    q = _queries(V, 20000, 2)
    sc = S.Scene(V, F).build_bvh()
    gi, gd = (t.cpu().numpy() for t in sc.closest_point(q))
    o = oracle.build_tree(V, F)
    oi, od = oracle.closest_point(o, q)
    sq = lambda tri, qq: oracle.tri_dist(V, F, tri, qq) ** 2  # noqa: E731
    _check_ties(oracle, V, F, q, gi.view(np.uint32), oi, gd, od, sq)


@pytest.mark.parametrize("name", ["ico5", "torus", "soup", "one", "two", "three"])
@pytest.mark.parametrize("flip", [False, True])
@pytest.mark.parametrize("reference_radius", [False, True])
def test_closest_silhouette_parity(S, oracle, name, flip, reference_radius):
    """reference_radius: the reference's midpoint cone radius, with which the
    traversal uses the reference's acos-based cone test instead of the
    trig-free one; GPU and oracle must still agree."""
    V, F = MESHES[name]()
    q = _queries(V, 20000, 3)
    sc = S.Scene(V, F).build_bvh(reference_radius=reference_radius)
    ge, gd = (t.cpu().numpy() for t in sc.closest_silhouette(q, flip=flip))
    o = oracle.build_tree(V, F, conservative_radius=not reference_radius)
    oe, od = oracle.closest_silhouette(o, q, flip=flip)
    f = lambda e, qq: oracle.edge_sil_dist(V, o.slots, e, qq, flip=flip)  # noqa: E731
    _check_ties(oracle, V, F, q, ge.view(np.uint32), oe, gd, od, f)


@pytest.mark.parametrize("name", ["ico5", "torus", "soup", "two"])
def test_ray_parity(S, oracle, name):
    V, F = MESHES[name]()
    n = 20000
    o_ = _queries(V, n, 4)
    d = W.unit_directions(n, 5)
    tmax = np.where(np.arange(n) % 3 == 0, np.float32(0.3), np.float32(np.inf)).astype(np.float32)
    sc = S.Scene(V, F).build_bvh()
    hit, t, uv, idx = (x.cpu().numpy() for x in sc.ray_intersect(o_, d, tmax))
    tr = oracle.build_tree(V, F)
    ohit, ot, ouv, oidx = oracle.ray(tr, o_, d, tmax)
    np.testing.assert_array_equal(hit, ohit)
    np.testing.assert_array_equal(t[hit], ot[hit])
    same = idx.view(np.uint32) == oidx
    np.testing.assert_array_equal(uv[hit & same], ouv[hit & same])
    ahit = sc.ray_intersect(o_, d, tmax, any_hit=True)[0].cpu().numpy()
    np.testing.assert_array_equal(ahit, ohit)


@pytest.mark.parametrize("name", ["ico5", "torus", "soup", "one", "two"])
def test_sample_bit_exact(S, oracle, name):
    V, F = MESHES[name]()
    n = 20000
    c = _queries(V, n, 6)
    r = W.log_uniform(n, 1e-3, 0.5, 7)
    u = np.random.default_rng(8).random((n, 3)).astype(np.float32)
    sc = S.Scene(V, F).build_bvh()
    gi, gp, gpt = (x.cpu().numpy() for x in sc.sample_in_sphere(c, r, u))
    tr = oracle.build_tree(V, F)
    oi, op, opt = oracle.sample_in_sphere(tr, c, r, u)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(gpt, opt)
    ok = oi >= 0
    pt2 = sc.sample_on_object(gi[ok], u[ok, 1:]).cpu().numpy()
    np.testing.assert_array_equal(pt2, opt[ok])


def test_seeded_silhouette_identical(S):
    V, F = W.cut_torus(96, 48)
    q = _queries(V, 30000, 9)
    sc = S.Scene(V, F).build_bvh()
    e0, d0 = (t.cpu().numpy() for t in sc.closest_silhouette(q))
    rng = np.random.default_rng(1)
    mixed = np.where(rng.random(len(q)) < 0.02, d0 * 0.5, d0 * 1.1)  # 2 % unsettled, scattered
    for seed in (np.zeros(len(q)), d0 * 0.5, d0, d0 * 1.0001 + 1e-3, rng.random(len(q)) * 0.3,
                 np.full(len(q), -1.0), np.full(len(q), np.nan), mixed, np.full(len(q), np.inf)):
        # unsettled queries re-queued for whole waves (default), or searched again in place
        for in_place in (False, True):
            e1, d1 = (t.cpu().numpy() for t in sc.closest_silhouette(q, seed_radius=seed.astype(np.float32),
                                                                     seed_in_place=in_place))
            np.testing.assert_array_equal(e1, e0)
            np.testing.assert_array_equal(d1, d0)
    # a query count that is not a multiple of the 64-query chunk, and a single query
    for m in (1, 63, 4097):
        e1, d1 = (t.cpu().numpy() for t in sc.closest_silhouette(q[:m], seed_radius=mixed[:m].astype(np.float32)))
        np.testing.assert_array_equal(e1, e0[:m])
        np.testing.assert_array_equal(d1, d0[:m])
    for bs in (64, 128, 256, 512):
        for depth in (8, 16, 24, 32):
            e1, d1 = (t.cpu().numpy() for t in sc.closest_silhouette(q, block_size=bs, stack_depth=depth))
            np.testing.assert_array_equal(e1, e0)
            np.testing.assert_array_equal(d1, d0)
    # a hard cap: exact within the radius, beyond it only "farther than the cap"
    i0, c0 = (t.cpu().numpy() for t in sc.closest_point(q))
    for bs in (64, 256):
        for depth in (8, 16):
            i1, c1 = (t.cpu().numpy() for t in sc.closest_point(q, block_size=bs, stack_depth=depth))
            np.testing.assert_array_equal(i1, i0)
            np.testing.assert_array_equal(c1, c0)
    for cap in (d0, d0 * 0.5, rng.random(len(q)).astype(np.float32) * 0.3, c0):
        cap = np.where(np.isfinite(cap), cap, np.float32(1.0)).astype(np.float32)
        e1, d1 = (t.cpu().numpy() for t in sc.closest_silhouette(q, seed_radius=cap, seed_is_cap=True))
        within = d0 <= cap
        np.testing.assert_array_equal(d1[within], d0[within])
        np.testing.assert_array_equal(e1[within], e0[within])
        assert np.all(d1[~within] > cap[~within])
        # the default shape (persistent launch) equals one block per chunk bit
        # for bit, beyond the cap too
        e2, d2 = (t.cpu().numpy() for t in sc.closest_silhouette(q, seed_radius=cap, seed_is_cap=True,
                                                                  block_size=128))
        np.testing.assert_array_equal(e2, e1)
        np.testing.assert_array_equal(d2, d1)


def test_visiting_order_invariance(S, oracle):
    """Morton visiting order (automatic from 65536 queries, forced, or
    supplied by the caller) gives the input-order results bit for bit, and
    those match the oracle."""
    import torch

    V, F = W.cut_torus(96, 48)
    q = _queries(V, 70000, 12)
    sc = S.Scene(V, F).build_bvh()
    perm = sc.query_order(q).cpu().numpy()
    assert np.array_equal(np.sort(perm), np.arange(len(q)))
    ref_cp = [t.cpu().numpy() for t in sc.closest_point(q, coherent=2)]
    ref_sil = [t.cpu().numpy() for t in sc.closest_silhouette(q, coherent=2)]
    order = torch.from_numpy(perm[::-1].copy()).cuda()
    for kw in ({}, {"coherent": 1}, {"order": order}):
        for got, ref in zip(sc.closest_point(q, **kw), ref_cp):
            np.testing.assert_array_equal(got.cpu().numpy(), ref)
        for got, ref in zip(sc.closest_silhouette(q, **kw), ref_sil):
            np.testing.assert_array_equal(got.cpu().numpy(), ref)
    o = oracle.build_tree(V, F)
    _, od = oracle.closest_silhouette(o, q[:8192])
    np.testing.assert_array_equal(ref_sil[1][:8192], od)


def test_silhouette_vs_fp64_bruteforce(S, oracle):
    V, F = W.cut_torus(64, 32)
    q = _queries(V, 4096, 10)
    sc = S.Scene(V, F).build_bvh()
    ge, gd = (t.cpu().numpy() for t in sc.closest_silhouette(q))
    slots, _ = sc.silhouettes()
    be, bd = oracle.bf_closest_silhouette(V, slots, q)
    fin = np.isfinite(bd)
    assert np.array_equal(np.isfinite(gd), fin)
    d64 = oracle.edge_sil_dist64(V, slots, ge.view(np.uint32)[fin], q[fin])
    assert np.all(d64 - bd[fin] <= 1e-6)


def test_error_paths(S):
    import ctypes

    L = S.lib()
    V, F = MESHES["two"]()
    with pytest.raises(S.SnchError) as e:
        S.Scene(V, np.zeros((0, 3), np.int32))
    assert e.value.code == 2
    bad = V.copy()
    bad[0, 0] = np.nan
    with pytest.raises(S.SnchError) as e:
        S.Scene(bad, F)
    assert e.value.code == 1
    with pytest.raises(S.SnchError) as e:
        S.Scene(V, np.int32([[0, 1, 7]]))
    assert e.value.code == 1
    sc = S.Scene(V, F)
    with pytest.raises(S.SnchError) as e:
        sc.closest_point(np.zeros((4, 3), np.float32))
    assert e.value.code == 4 and "not built" in str(e.value)
    sc.build_bvh()
    idx, d = sc.closest_point(np.zeros((0, 3), np.float32))
    assert idx.numel() == 0
    assert L.snch_closest_point(sc._h, None, 5, None, None, None) == 1


def test_companion_row_counts(S):
    """Every per-query input must hold one row per query; a short array is
    rejected before any launch."""
    V, F = MESHES["ico5"]()
    sc = S.Scene(V, F).build_bvh()
    q = _queries(V, 100, 2)
    with pytest.raises(ValueError):
        sc.closest_silhouette(q, seed_radius=np.ones(99, np.float32))
    with pytest.raises(ValueError):
        sc.ray_intersect(q, W.unit_directions(99, 5))
    with pytest.raises(ValueError):
        sc.ray_intersect(q, W.unit_directions(100, 5), tmax=np.ones(50, np.float32))
    with pytest.raises(ValueError):
        sc.sample_in_sphere(q, np.ones(99, np.float32), np.zeros((100, 3), np.float32))
    with pytest.raises(ValueError):
        sc.sample_in_sphere(q, np.ones(100, np.float32), np.zeros((10, 3), np.float32))
    with pytest.raises(ValueError):
        sc.sample_on_object(np.zeros(5, np.int32), np.zeros((4, 2), np.float32))


def test_two_streams_share_scratch_safely(S):
    """Ordered queries (which use the handle's ordering scratch) issued on two
    streams back to back give the results of serial calls."""
    import torch

    V, F = W.cut_torus(256, 128)
    sc = S.Scene(V, F).build_bvh()
    qa = torch.from_numpy(_queries(V, 300000, 41)).cuda()
    qb = torch.from_numpy(_queries(V, 250000, 42)).cuda()
    ra = [t.clone() for t in sc.closest_silhouette(qa)]
    rb = [t.clone() for t in sc.closest_silhouette(qb)]
    ca = [t.clone() for t in sc.closest_point(qa)]
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):
        with torch.cuda.stream(s1):
            ga = sc.closest_silhouette(qa)
        with torch.cuda.stream(s2):
            gb = sc.closest_silhouette(qb)
        with torch.cuda.stream(s1):
            gc = sc.closest_point(qa)
        torch.cuda.synchronize()
        for got, ref in zip(ga + gb + gc, ra + rb + ca):
            assert torch.equal(got, ref)
