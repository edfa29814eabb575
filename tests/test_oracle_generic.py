"""Pins the oracle's generic overlap and line queries (query.cuh:12-77,
:176-231): the id sets equal a brute force over the leaves, and the emission
order equals an independent recursive restatement of the reference's stack
order (pop a node; left child, then right; emit leaves, push internal
children)."""
import numpy as np
import pytest

from snch_lbvh_amd import workloads as W


def _stack_order(tree, hit_box, N):
    """The reference's order, restated with an explicit Python stack."""
    if N == 1:
        return [int(tree.nodes[0, 3])] if hit_box(0) else []
    out, stack = [], [0]
    while stack:
        node = stack.pop()
        for c in (int(tree.nodes[node, 1]), int(tree.nodes[node, 2])):
            if hit_box(c):
                o = int(tree.nodes[c, 3])
                if o != 0xFFFFFFFF:
                    out.append(o)
                else:
                    stack.append(c)
    return out


@pytest.mark.parametrize("which", ["ico", "torus", "one"])
def test_overlap_oracle(oracle, which):
    if which == "ico":
        V, F = W.icosphere(3, 0.02, seed=1)
    elif which == "torus":
        V, F = W.cut_torus(32, 16)
    else:
        V, F = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]]), np.int32([[0, 1, 2]])
    t = oracle.build_tree(V, F)
    N = len(F)
    rng = np.random.default_rng(2)
    lo0, hi0 = W.padded_box(V, 0.2)
    c = rng.uniform(lo0, hi0, (300, 3)).astype(np.float32)
    e = (rng.random((300, 3)) * 0.3).astype(np.float32)
    lo, hi = c - e, c + e
    cnt, idx = oracle.overlap(t, lo, hi, 4096)
    leaf_boxes = t.boxes[N - 1:] if N > 1 else t.boxes
    leaf_obj = t.nodes[N - 1:, 3] if N > 1 else t.nodes[:, 3]
    for i in range(len(lo)):
        hitb = lambda k: bool(np.all(t.boxes[k, 3:] >= lo[i]) and np.all(hi[i] >= t.boxes[k, :3]))  # noqa: E731
        brute = {int(o) for b, o in zip(leaf_boxes, leaf_obj)
                 if np.all(b[3:] >= lo[i]) and np.all(hi[i] >= b[:3])}
        got = list(idx[i, :cnt[i]])
        assert set(got) == brute and len(got) == len(brute)
        assert got == _stack_order(t, hitb, N)
    # truncation keeps the first max_out in order and the full count
    cnt2, idx2 = oracle.overlap(t, lo, hi, 3)
    np.testing.assert_array_equal(cnt2, cnt)
    for i in range(len(lo)):
        k = min(3, cnt[i])
        np.testing.assert_array_equal(idx2[i, :k], idx[i, :k])


def _line_box32(b, o, inv):
    t1 = (b[0] - o[0]) * inv[0]
    t2 = (b[3] - o[0]) * inv[0]
    tmin, tmax = np.fmin(t1, t2), np.fmax(t1, t2)
    for k in (1, 2):
        t1 = (b[k] - o[k]) * inv[k]
        t2 = (b[3 + k] - o[k]) * inv[k]
        tmin = np.fmax(tmin, np.fmin(t1, t2))
        tmax = np.fmin(tmax, np.fmax(t1, t2))
    return bool(tmax >= tmin)


def _line_tri64(o, d, a, b, c):
    """FP64 Moller-Trumbore on the infinite line: (u, v, t) or None."""
    o, d, a, b, c = (np.asarray(x, np.float64) for x in (o, d, a, b, c))
    e1, e2 = b - a, c - a
    h = np.cross(d, e2)
    det = e1 @ h
    if abs(det) < 1e-12:
        return None
    s = o - a
    u = (s @ h) / det
    q = np.cross(s, e1)
    v = (d @ q) / det
    return u, v, (e2 @ q) / det


def test_line_oracle(oracle):
    V, F = W.icosphere(3, 0.02, seed=1)
    t = oracle.build_tree(V, F)
    N = len(F)
    rng = np.random.default_rng(3)
    n = 120
    o = rng.uniform(-1.3, 1.3, (n, 3)).astype(np.float32)
    d = W.unit_directions(n, 4)
    cnt, idx, tt = oracle.line_intersect(t, o, d, 64)
    closed = 0
    for i in range(n):
        inv = np.float32(1.0) / d[i]
        got = list(idx[i, :cnt[i]])
        # the box-hit leaves in stack order, filtered by the element test
        order = _stack_order(t, lambda k: _line_box32(t.boxes[k], o[i], inv), N)
        keep = set(int(x) for x in got)
        assert keep <= set(order)
        assert got == [x for x in order if x in keep]
        for j, tri in enumerate(got):
            r = _line_tri64(o[i], d[i], *V[F[tri]])
            assert r is not None
            u, v, tl = r
            assert u >= -1e-5 and v >= -1e-5 and u + v <= 1 + 1e-5
            assert abs(tt[i, j] - tl) <= 1e-5 * (1 + abs(tl))
        # every clear crossing (away from the triangle's edges) is found
        for tri in range(N):
            r = _line_tri64(o[i], d[i], *V[F[tri]])
            if r is not None and r[0] > 1e-4 and r[1] > 1e-4 and r[0] + r[1] < 1 - 1e-4:
                assert tri in got
        # a line through a closed surface crosses it an even number of times
        closed += cnt[i] % 2 == 0
    assert closed >= n - 2


def test_line_oracle_2d(oracle):
    V, S = W.circle_in_square(64, 16)
    t = oracle.build_tree2d(V, S)
    rng = np.random.default_rng(4)
    n = 300
    o = rng.uniform(-0.9, 0.9, (n, 2)).astype(np.float32)
    ang = rng.uniform(0, 2 * np.pi, n)
    d = np.stack([np.cos(ang), np.sin(ang)], 1).astype(np.float32)
    cnt, idx, tt = oracle.line_intersect(t, o, d, 64)
    on_square = np.all(np.abs(V[S]).max(2) > 0.99, axis=1)  # both end points on the outer square
    # a line crosses the outer square exactly twice (more only through its corners)
    for i in range(n):
        got = idx[i, :cnt[i]]
        sq = int(np.sum(on_square[got]))
        assert 2 <= sq <= 4, (i, got)
        p = o[i].astype(np.float64) + tt[i, :cnt[i], None] * d[i]
        seg = V[S[got]].astype(np.float64)
        # the points lie on their segments' lines (parameter within the reference's 1e-3 tolerance)
        dirs = seg[:, 1] - seg[:, 0]
        cross = dirs[:, 0] * (p[:, 1] - seg[:, 0, 1]) - dirs[:, 1] * (p[:, 0] - seg[:, 0, 0])
        assert np.all(np.abs(cross) < 1e-4)
