"""GPU parity for scene<2> (segments, silhouette vertices) against the oracle."""
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
    "cis": lambda: W.circle_in_square(400, 50, sigma=0.01),
    "soup": lambda: W.polyline_soup(300, 500),
    "one": lambda: (np.float32([[0, 0], [1, 0]]), np.int32([[0, 1]])),
    "two": lambda: (np.float32([[0, 0], [1, 0], [1, 1]]), np.int32([[0, 1], [1, 2]])),
}


def _q(n, seed):
    return np.random.default_rng(seed).uniform(-1.2, 1.2, (n, 2)).astype(np.float32)


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("reference_radius", [False, True])
def test_build_2d_bit_exact(S, oracle, name, reference_radius):
    V, Sg = MESHES[name]()
    sc = S.Scene2D(V, Sg).build_bvh(reference_radius=reference_radius)
    g = sc.export()
    o = oracle.build_tree2d(V, Sg, conservative_radius=not reference_radius)
    slots, owned = sc.silhouettes()
    np.testing.assert_array_equal(slots, o.slots)
    np.testing.assert_array_equal(owned, o.owned)
    for k, ov in (("sorted_codes", o.sorted_codes), ("sorted_idx", o.sorted_idx), ("nodes", o.nodes),
                  ("boxes", o.boxes), ("cones", o.cones)):
        np.testing.assert_array_equal(g[k], ov, err_msg=k)


@pytest.mark.parametrize("name", list(MESHES))
def test_queries_2d(S, oracle, name):
    V, Sg = MESHES[name]()
    sc = S.Scene2D(V, Sg).build_bvh()
    o = oracle.build_tree2d(V, Sg)
    q = _q(20000, 3)
    gi, gd = (t.cpu().numpy() for t in sc.closest_point(q))
    oi, od = oracle.closest_point2d(o, q)
    np.testing.assert_array_equal(gd, od)
    for flip in (False, True):
        ge, gsd = (t.cpu().numpy() for t in sc.closest_silhouette(q, flip=flip))
        oe, osd = oracle.closest_silhouette2d(o, q, flip=flip)
        np.testing.assert_array_equal(gsd, osd)
        diff = ge.view(np.uint32) != oe
        # a different id at an identical distance can only be a tie at the same vertex position
        assert np.all(np.isin(np.where(diff)[0], np.where(np.isfinite(osd))[0]))
    d = np.random.default_rng(4).normal(size=(20000, 2)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    hit, t, s, idx = (x.cpu().numpy() for x in sc.ray_intersect(q, d))
    ohit, ot, os_, oidx = oracle.ray2d(o, q, d)
    np.testing.assert_array_equal(hit, ohit)
    np.testing.assert_array_equal(t[hit], ot[hit])
    r = np.random.default_rng(5).uniform(0.01, 0.5, 20000).astype(np.float32)
    u = np.random.default_rng(6).random((20000, 3)).astype(np.float32)
    si, sp, spt = (x.cpu().numpy() for x in sc.sample_in_sphere(q, r, u))
    oi2, op2, opt2 = oracle.sample2d(o, q, r, u)
    np.testing.assert_array_equal(si, oi2)
    np.testing.assert_array_equal(sp, op2)
    np.testing.assert_array_equal(spt, opt2)
