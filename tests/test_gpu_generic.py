"""GPU overlap and line queries (query.cuh:12-77, :176-231) against the
oracle: counts, ids in the reference's visit order, and line parameters, bit
for bit, including truncation at max_per_query and counts-only calls."""
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


MESHES = {
    "ico5": lambda: W.icosphere(5, 0.02, seed=1),
    "torus": lambda: W.cut_torus(96, 48),
    "soup": lambda: (np.random.default_rng(9).random((300, 3)).astype(np.float32),
                     np.random.default_rng(10).integers(0, 300, size=(500, 3)).astype(np.int32)),
    "one": lambda: (np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]]), np.int32([[0, 1, 2]])),
    "two": lambda: (np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.5]]), np.int32([[0, 1, 2], [1, 3, 2]])),
}


def _boxes(V, n, seed, dim=3):
    rng = np.random.default_rng(seed)
    lo0, hi0 = V.min(0), V.max(0)
    ext = hi0 - lo0
    c = rng.uniform(lo0 - 0.1 * ext, hi0 + 0.1 * ext, (n, dim)).astype(np.float32)
    e = (rng.random((n, dim)) * 0.15 * ext.max()).astype(np.float32)
    return c - e, c + e


@pytest.mark.parametrize("name", list(MESHES))
def test_overlap_parity(S, oracle, name):
    V, F = MESHES[name]()
    sc = S.Scene(V, F).build_bvh()
    t = oracle.build_tree(V, F)
    lo, hi = _boxes(V, 5000, 1)
    for mx in (256, 8, 0):
        gc, gi = (x.cpu().numpy() for x in sc.overlap(lo, hi, mx))
        oc, oi = oracle.overlap(t, lo, hi, mx)
        np.testing.assert_array_equal(gc.view(np.uint32), oc)
        if mx:
            k = np.minimum(oc, mx)
            mask = np.arange(mx)[None, :] < k[:, None]
            np.testing.assert_array_equal(gi.view(np.uint32)[mask], oi[mask])
            assert np.all(gi[~mask] == -1)


@pytest.mark.parametrize("name", list(MESHES))
def test_line_parity(S, oracle, name):
    V, F = MESHES[name]()
    sc = S.Scene(V, F).build_bvh()
    t = oracle.build_tree(V, F)
    lo, hi = W.padded_box(V, 0.1)
    o = W.uniform_points(5000, lo, hi, 3)
    d = W.unit_directions(5000, 4)
    for mx in (64, 4):
        gc, gi, gt = (x.cpu().numpy() for x in sc.line_intersect(o, d, mx))
        oc, oi, ot = oracle.line_intersect(t, o, d, mx)
        np.testing.assert_array_equal(gc.view(np.uint32), oc)
        k = np.minimum(oc, mx)
        mask = np.arange(mx)[None, :] < k[:, None]
        np.testing.assert_array_equal(gi.view(np.uint32)[mask], oi[mask])
        np.testing.assert_array_equal(gt[mask], ot[mask])


def test_generic_2d(S, oracle):
    V, Sg = W.circle_in_square(200, 30, sigma=0.02)
    sc = S.Scene2D(V, Sg).build_bvh()
    t = oracle.build_tree2d(V, Sg)
    lo, hi = _boxes(V, 4000, 5, dim=2)
    gc, gi = (x.cpu().numpy() for x in sc.overlap(lo, hi, 64))
    oc, oi = oracle.overlap(t, lo, hi, 64)
    np.testing.assert_array_equal(gc.view(np.uint32), oc)
    mask = np.arange(64)[None, :] < np.minimum(oc, 64)[:, None]
    np.testing.assert_array_equal(gi.view(np.uint32)[mask], oi[mask])
    rng = np.random.default_rng(6)
    o = rng.uniform(-1, 1, (4000, 2)).astype(np.float32)
    ang = rng.uniform(0, 2 * np.pi, 4000)
    d = np.stack([np.cos(ang), np.sin(ang)], 1).astype(np.float32)
    gc, gi, gt = (x.cpu().numpy() for x in sc.line_intersect(o, d, 32))
    oc, oi, ot = oracle.line_intersect(t, o, d, 32)
    np.testing.assert_array_equal(gc.view(np.uint32), oc)
    mask = np.arange(32)[None, :] < np.minimum(oc, 32)[:, None]
    np.testing.assert_array_equal(gi.view(np.uint32)[mask], oi[mask])
    np.testing.assert_array_equal(gt[mask], ot[mask])


def test_generic_large(S, oracle):
    """The 2^20-triangle torus: 2^20 small boxes and 2^18 lines against a
    16,384-query oracle subsample; all counts are consistent with a rerun."""
    import torch

    V, F = W.cut_torus(1024, 512, seed=3)
    sc = S.Scene(V, F).build_bvh()
    t = oracle.build_tree(V, F)
    lo, hi = _boxes(V, 1 << 20, 7)
    hi = lo + (hi - lo) * 0.05
    gc, gi = sc.overlap(lo, hi, 32)
    sub = np.sort(np.random.default_rng(8).choice(1 << 20, 16384, replace=False))
    oc, oi = oracle.overlap(t, lo[sub], hi[sub], 32)
    np.testing.assert_array_equal(gc.cpu().numpy().view(np.uint32)[sub], oc)
    m = np.arange(32)[None, :] < np.minimum(oc, 32)[:, None]
    np.testing.assert_array_equal(gi.cpu().numpy().view(np.uint32)[sub][m], oi[m])
    gc2, _ = sc.overlap(lo, hi, 0)
    assert torch.equal(gc, gc2)
    blo, bhi = W.padded_box(V)
    o = W.uniform_points(1 << 18, blo, bhi, 9)
    d = W.unit_directions(1 << 18, 10)
    lc, li, lt = (x.cpu().numpy() for x in sc.line_intersect(o, d, 16))
    sub = sub[sub < (1 << 18)]
    oc, oi, ot = oracle.line_intersect(t, o[sub], d[sub], 16)
    np.testing.assert_array_equal(lc.view(np.uint32)[sub], oc)
    m = np.arange(16)[None, :] < np.minimum(oc, 16)[:, None]
    np.testing.assert_array_equal(li.view(np.uint32)[sub][m], oi[m])
    np.testing.assert_array_equal(lt[sub][m], ot[m])
