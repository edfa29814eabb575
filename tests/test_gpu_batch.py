"""Batched meshes (one build, one persistent query launch) and leaf runs.

Bars: every mesh's tree in a batch equals the oracle's tree of that mesh
built alone, bit for bit; batched query results equal the single-scene
results bit for bit, for every launch configuration the autotuner can pick;
leaf size 2, 4, 8 leaves every query result unchanged."""
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


def _soup(seed=9, nv=200, nf=300):
    rng = np.random.default_rng(seed)
    return rng.random((nv, 3)).astype(np.float32) + 2.0, rng.integers(0, nv, size=(nf, 3)).astype(np.int32)


def _meshes():
    return [
        W.icosphere(3, 0.02, seed=5),            # 1280
        W.cut_torus(64, 40, seed=6),              # 5120: several refit subtrees
        (np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]]), np.int32([[0, 1, 2]])),
        _soup(),
        W.icosphere(2, 0.0),                      # 320
        (np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.5]]), np.int32([[0, 1, 2], [1, 3, 2]])),
        W.uv_sphere(3000, seed=7),
    ]


def _queries(meshes, counts, seed):
    qs = []
    for (V, _), n in zip(meshes, counts):
        lo, hi = W.padded_box(V, 0.25)
        qs.append(W.uniform_points(n, lo, hi, seed + len(qs)))
    q_off = np.zeros(len(qs) + 1, np.uint32)
    q_off[1:] = np.cumsum(counts)
    return np.concatenate(qs).astype(np.float32), q_off


def test_batch_build_matches_oracle(S, oracle):
    meshes = _meshes()
    b = S.BatchScene(meshes).build_bvh()
    for m, (V, F) in enumerate(meshes):
        g = b.export(m)
        o = oracle.build_tree(V, F)
        for k, ov in (("sorted_codes", o.sorted_codes), ("sorted_idx", o.sorted_idx), ("nodes", o.nodes),
                      ("boxes", o.boxes), ("cones", o.cones)):
            np.testing.assert_array_equal(g[k], ov, err_msg=f"mesh {m} {k}")


def test_batch_build_refit_levels(S, oracle):
    """Meshes of every refit class in one batch: one refit subtree (at most
    1,024 triangles), a mid set only (1,025 to 16,384), mid sets and nodes
    above them (more); each mesh's tree equals the oracle's with every layout
    of the pass above the mid sets."""
    meshes = [
        W.cut_torus(200, 100, seed=11),           # 40,000
        W.icosphere(2, 0.0),                      # 320
        W.cut_torus(96, 48, seed=12),             # 9,216
        W.icosphere(5, 0.02, seed=13),            # 20,480
        (np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]]), np.int32([[0, 1, 2]])),
        W.uv_sphere(17000, seed=14),
    ]
    orc = [oracle.build_tree(V, F) for V, F in meshes]
    b = S.BatchScene(meshes)
    for layout in ("auto", "memory", "climb"):
        b.build_bvh(top_layout=layout)
        for m, o in enumerate(orc):
            g = b.export(m)
            for k, ov in (("sorted_idx", o.sorted_idx), ("nodes", o.nodes), ("boxes", o.boxes), ("cones", o.cones)):
                np.testing.assert_array_equal(g[k], ov, err_msg=f"{layout}: mesh {m} {k}")


@pytest.mark.parametrize("tune", [None, {"block_size": 64, "stack_depth": 24, "queries_per_wave": 128},
                                  {"block_size": 512, "stack_depth": 32, "queries_per_wave": 256},
                                  {"block_size": 128, "stack_depth": 16, "queries_per_wave": 64}])
def test_batch_queries_match_single(S, tune):
    meshes = _meshes()
    counts = [3000, 5000, 70, 0, 900, 33, 4000]
    q, q_off = _queries(meshes, counts, 20)
    b = S.BatchScene(meshes).build_bvh()
    singles = [S.Scene(V, F).build_bvh() for V, F in meshes]
    for coherent in (1, 2):
        bi, bd = (t.cpu().numpy() for t in b.closest_point(q, q_off, tune=tune, coherent=coherent))
        si, sd = (t.cpu().numpy() for t in b.closest_silhouette(q, q_off, tune=tune, coherent=coherent))
        for m, sc in enumerate(singles):
            a, z = q_off[m], q_off[m + 1]
            if a == z:
                continue
            ri, rd = (t.cpu().numpy() for t in sc.closest_point(q[a:z]))
            np.testing.assert_array_equal(bi[a:z], ri, err_msg=f"cp idx mesh {m}")
            np.testing.assert_array_equal(bd[a:z], rd, err_msg=f"cp dist mesh {m}")
            re, rs = (t.cpu().numpy() for t in sc.closest_silhouette(q[a:z]))
            np.testing.assert_array_equal(si[a:z], re, err_msg=f"sil id mesh {m}")
            np.testing.assert_array_equal(sd[a:z], rs, err_msg=f"sil dist mesh {m}")


@pytest.mark.parametrize("leaf_size", [2, 4, 8])
def test_leaf_size_results_unchanged(S, leaf_size):
    V, F = W.cut_torus(96, 48)
    lo, hi = W.padded_box(V, 0.25)
    q = W.uniform_points(20000, lo, hi, 31)
    d = W.unit_directions(20000, 32)
    a = S.Scene(V, F).build_bvh()
    b = S.Scene(V, F).build_bvh(leaf_size=leaf_size)
    for x, y in zip(a.closest_point(q), b.closest_point(q)):
        np.testing.assert_array_equal(x.cpu().numpy(), y.cpu().numpy())
    for flip in (False, True):
        for x, y in zip(a.closest_silhouette(q, flip=flip), b.closest_silhouette(q, flip=flip)):
            np.testing.assert_array_equal(x.cpu().numpy(), y.cpu().numpy())
    for x, y in zip(a.ray_intersect(q, d), b.ray_intersect(q, d)):
        np.testing.assert_array_equal(x.cpu().numpy(), y.cpu().numpy())
    # the exported tree (reference layout) does not depend on the leaf size
    ga, gb = a.export(), b.export()
    for k in ga:
        np.testing.assert_array_equal(ga[k], gb[k])


def test_autotune_then_query(S):
    meshes = _meshes()[:3] + [W.cut_torus(48, 24, seed=8)]
    counts = [2000, 3000, 100, 2500]
    q, q_off = _queries(meshes, counts, 40)
    b = S.BatchScene(meshes).build_bvh()
    ref = [t.cpu().numpy() for t in b.closest_silhouette(q, q_off)]
    r = b.autotune(q, q_off, kind="silhouette", sample_per_mesh=0)
    assert r["leaf_size"] in (1, 2, 4, 8) and r["block_size"] in (64, 128, 256, 512)
    assert r["stack_depth"] in (8, 16, 24, 32) and r["queries_per_wave"] in (64, 128, 256)
    assert r["n_tried"] == 1 + 3 + 3 + 3 + 2 and 0 < r["ms"] <= r["ms_default"]
    got = [t.cpu().numpy() for t in b.closest_silhouette(q, q_off, tune=r)]
    for x, y in zip(got, ref):
        np.testing.assert_array_equal(x, y)


def test_batch_errors(S):
    meshes = _meshes()[:2]
    b = S.BatchScene(meshes)
    q, q_off = _queries(meshes, [10, 10], 50)
    with pytest.raises(S.SnchError) as e:
        b.closest_point(q, q_off)
    assert e.value.code == 4
    b.build_bvh()
    with pytest.raises(S.SnchError):
        b.closest_point(q, q_off, tune={"block_size": 96})
    with pytest.raises(S.SnchError):
        b.closest_point(q, q_off, tune={"queries_per_wave": 100})
    with pytest.raises(S.SnchError):
        S.BatchScene([meshes[0], (np.zeros((3, 3), np.float32), np.zeros((0, 3), np.int32))])
    # q_off: n_meshes + 1 non-decreasing offsets ending at the query count
    with pytest.raises(ValueError):
        b.closest_point(q, q_off[:-1])
    with pytest.raises(ValueError):
        b.closest_silhouette(q, np.array([0, 15, 5], np.uint32))
    with pytest.raises(ValueError):
        b.closest_silhouette(q, np.array([0, 10, 30], np.uint32))
    with pytest.raises(ValueError):
        b.autotune(q, np.array([1, 10, 20], np.uint32))


def test_batch_seeded_identical(S):
    """Seeded batched silhouette queries (exact, loose, too small, none, NaN)
    equal the unseeded ones."""
    import torch

    meshes = _meshes()
    b = S.BatchScene(meshes).build_bvh()
    counts = [3000, 500, 70, 900, 130, 10, 2000]
    q, q_off = _queries(meshes, counts, 77)
    e0, d0 = b.closest_silhouette(q, q_off)
    rng = np.random.default_rng(3)
    d = d0.cpu().numpy()
    fin = np.isfinite(d)
    for seed in (np.where(fin, d, 1.0), np.where(fin, d * 1.05, 0.1), np.zeros(len(q)), rng.random(len(q)) * 0.5,
                 np.full(len(q), np.nan)):
        e1, d1 = b.closest_silhouette(q, q_off, seed_radius=seed.astype(np.float32))
        assert torch.equal(e1, e0) and torch.equal(d1, d0)
        e2, d2 = b.closest_silhouette(q, q_off, seed_radius=seed.astype(np.float32), tune={"block_size": 256})
        assert torch.equal(e2, e0) and torch.equal(d2, d0)
