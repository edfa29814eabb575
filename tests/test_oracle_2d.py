"""Pins the 2D (scene<2>) part of the oracle: silhouette table against a
literal restatement, topology, refit closed form, FP64 brute force, KATs."""
import numpy as np
import pytest

from snch_lbvh_amd import workloads as W


def _literal_silhouettes(S, nV):
    slots = [[-1, -1, -1, -1] for _ in range(nV)]
    for a, b in S:
        slots[a][1], slots[a][2] = int(a), int(b)
        slots[b][0], slots[b][1] = int(a), int(b)
    seen, owned = set(), []
    for a, b in S:
        o = []
        for v in (int(a), int(b)):
            if v not in seen:
                seen.add(v)
                o.append(v)
        owned.append(o + [-1] * (2 - len(o)))
    return np.array(slots), np.array(owned)


@pytest.mark.parametrize("which", ["cis", "soup"])
def test_silhouette_table_2d(oracle, which):
    V, S = W.circle_in_square(64, 16) if which == "cis" else W.polyline_soup(50, 90)
    slots, owned = oracle.silhouettes2d(S, V.shape[0])
    s2, o2 = _literal_silhouettes(S, V.shape[0])
    np.testing.assert_array_equal(slots, s2)
    np.testing.assert_array_equal(owned, o2)


def test_topology_and_refit_2d(oracle):
    V, S = W.circle_in_square(300, 40, sigma=0.01)
    t = oracle.build_tree2d(V, S)
    N = t.n_obj
    from test_oracle_core import _leaf_ranges, _topdown

    keys = (t.sorted_codes.astype(np.uint64) << np.uint64(32)) | t.sorted_idx.astype(np.uint64)
    left, right = _topdown([int(k) for k in keys])
    np.testing.assert_array_equal(t.nodes[: N - 1, 1], left)
    np.testing.assert_array_equal(t.nodes[: N - 1, 2], right)
    rr = _leaf_ranges(t)
    leaves = t.boxes[N - 1:]
    for n in range(N - 1):
        a, b = rr[n]
        np.testing.assert_array_equal(t.boxes[n, :3], leaves[a:b + 1, :3].min(0))
        np.testing.assert_array_equal(t.boxes[n, 3:], leaves[a:b + 1, 3:].max(0))
    assert np.all(t.boxes[:, 2] == 0) and np.all(t.cones[:, 2] == 0)


@pytest.mark.parametrize("which", ["cis", "soup"])
def test_queries_2d_vs_bruteforce(oracle, which):
    V, S = W.circle_in_square(200, 30, sigma=0.02) if which == "cis" else W.polyline_soup(300, 500)
    t = oracle.build_tree2d(V, S)
    q = np.random.default_rng(3).uniform(-1.2, 1.2, (4000, 2)).astype(np.float32)
    idx, d = oracle.closest_point2d(t, q)
    bi, bd = oracle.bf_closest_point2d(V, S, q)
    assert np.all(np.abs(d - bd) <= 1e-6 + 1e-6 * bd)
    sil, sd = oracle.closest_silhouette2d(t, q)
    bs, bsd = oracle.bf_closest_silhouette2d(V, t.slots, q)
    fin = np.isfinite(bsd)
    assert np.array_equal(np.isfinite(sd), fin)
    assert np.all(np.abs(sd[fin] - bsd[fin]) <= 1e-6 + 1e-6 * bsd[fin])


def test_kats_2d(oracle):
    V, S = W.circle_in_square(512, 64)
    t = oracle.build_tree2d(V, S)
    # points on the x axis between circle (r=0.5) and square: nearest boundary distance
    x = np.linspace(0.55, 0.95, 50).astype(np.float32)
    q = np.stack([x, np.zeros_like(x)], 1)
    _, d = oracle.closest_point2d(t, q)
    assert np.all(np.abs(d - np.minimum(x - 0.5, 1 - x)) < 2e-5)
    # rays from the origin along +x hit the circle at r ~ 0.5
    o = np.zeros((50, 2), np.float32)
    ang = np.linspace(0, 2 * np.pi, 50, endpoint=False)
    dd = np.stack([np.cos(ang), np.sin(ang)], 1).astype(np.float32)
    hit, tt, s, _ = oracle.ray2d(t, o, dd)
    assert hit.all() and np.all(np.abs(tt - 0.5) < 1e-3)
    # silhouette of a convex circle from outside: the tangent points, sqrt(rho^2 - r^2)
    q = np.float32([[0.0, 0.8], [0.7, 0.0]])
    _, sd = oracle.closest_silhouette2d(t, q)
    assert np.all(np.abs(sd - np.sqrt(np.sum(q.astype(np.float64) ** 2, 1) - 0.25)) < 1e-2)
    # sampler returns segments whose pdf is path probability / length
    c = np.zeros((1000, 2), np.float32) + np.float32([0.45, 0.0])
    r = np.full(1000, 0.2, np.float32)
    u = np.random.default_rng(1).random((1000, 3)).astype(np.float32)
    idx, pdf, pt = oracle.sample2d(t, c, r, u)
    ok = idx >= 0
    assert ok.mean() > 0.5 and np.all(pdf[ok] > 0)
    seg = V[S[idx[ok]]]
    # points lie on their segments
    cr = (seg[:, 1, 0] - seg[:, 0, 0]) * (pt[ok, 1] - seg[:, 0, 1]) - (seg[:, 1, 1] - seg[:, 0, 1]) * (pt[ok, 0] - seg[:, 0, 0])
    assert np.max(np.abs(cr)) < 1e-5
