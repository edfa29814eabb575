"""Pins for the oracle functions that the bit-exact comparisons alone leave
open: leaf normal cones against an FP64 restatement of scene.cuh:887-961,
the traversal's trig-free cone test against the reference's acos-based test,
the 2D sampler against exact path enumeration, the 2D ray against an FP64
brute force, and the walk-on-stars restatement against the exact solution of
a mixed Dirichlet / Neumann problem."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from snch_lbvh_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ------------------------------------------------------------ leaf cones
def _edge_normals64(V, s):
    """Unit face normals of an edge with slots (opp0, lo, hi, opp1) and their
    normalised sum (scene.cuh:746-787), in float64; a missing face gives a
    zero normal."""
    V = V.astype(np.float64)
    lo, hi = V[s[1]], V[s[2]]
    n = []
    if s[3] != -1:  # face 0: (hi - lo) x (opp1 - lo)
        n.append(np.cross(hi - lo, V[s[3]] - lo))
    else:
        n.append(None)
    if s[0] != -1:  # face 1: (lo - hi) x (opp0 - hi)
        n.append(np.cross(lo - hi, V[s[0]] - hi))
    else:
        n.append(None)
    raw = [x for x in n if x is not None]
    tot = np.sum(raw, axis=0)
    with np.errstate(divide="ignore", invalid="ignore"):  # a folded edge: 0 / 0, as in FP32
        unit = [x / np.linalg.norm(x) if x is not None else np.zeros(3) for x in n]
        return unit, tot / np.linalg.norm(tot), (s[0] != -1 and s[3] != -1)


def _folded(V, s):
    """An edge whose two faces nearly cancel (|n0 + n1| < 1e-3 (|n0| + |n1|)):
    its normal direction is FP32 rounding noise in the reference itself."""
    if s[0] == -1 or s[3] == -1:
        return False
    V = V.astype(np.float64)
    lo, hi = V[s[1]], V[s[2]]
    n0 = np.cross(hi - lo, V[s[3]] - lo)
    n1 = np.cross(lo - hi, V[s[0]] - hi)
    return np.linalg.norm(n0 + n1) < 1e-3 * (np.linalg.norm(n0) + np.linalg.norm(n1))


def _leaf_cones64(V, F, slots, owned, conservative):
    out = np.zeros((len(F), 5))
    eps = np.float32(1.1920928955078125e-07)
    for i, f in enumerate(F):
        P = V[f].astype(np.float32)
        blo = (P.min(0) - eps).astype(np.float64)
        bhi = (P.max(0) + eps).astype(np.float64)
        c = 0.5 * (blo + bhi)
        es = [e for e in owned[i] if e != -1]
        if not es:
            out[i] = [0, 0, 0, -np.pi, 0]
            continue
        axis = np.zeros(3)
        radius = 0.0
        two = True
        for e in es:
            s = slots[e]
            _, nsum, has2 = _edge_normals64(V, s)
            axis += nsum
            a, b = V[s[1]].astype(np.float64), V[s[2]].astype(np.float64)
            if conservative:
                radius = max(radius, np.linalg.norm(a - c), np.linalg.norm(b - c))
            else:
                radius = max(radius, np.linalg.norm(0.5 * (a + b) - c))
            two = two and has2
        if not two:
            out[i] = [*axis, np.pi, radius]
            continue
        an = np.linalg.norm(axis)
        if not an > 1.1920928955078125e-07:  # scene.cuh:930-933 (also NaN): the half angle stays pi
            out[i] = [*axis, np.pi, radius]
            continue
        axis = axis / an
        ha = 0.0
        for e in es:
            unit, _, _ = _edge_normals64(V, slots[e])
            for n in unit:
                ha = max(ha, np.arccos(np.clip(axis @ n, -1.0, 1.0)))
        out[i] = [*axis, ha, radius]
    return out


@pytest.mark.parametrize("which", ["ico", "torus", "soup"])
@pytest.mark.parametrize("conservative", [True, False])
def test_leaf_cones_fp64(oracle, which, conservative):
    if which == "ico":
        V, F = W.icosphere(3, 0.02, seed=3)
    elif which == "torus":
        V, F = W.cut_torus(40, 20, seed=4)  # boundary edges: half angle pi
    else:
        # a dense soup: most edges have two (or more: non-manifold) faces
        rng = np.random.default_rng(9)
        V = rng.random((25, 3)).astype(np.float32)
        F = rng.integers(0, 25, (300, 3)).astype(np.int32)
        F = F[(F[:, 0] != F[:, 1]) & (F[:, 1] != F[:, 2]) & (F[:, 0] != F[:, 2])]
    _, slots, owned = oracle.edge_table(F)
    _, cones = oracle.leaf_box_cone(V, F, conservative_radius=conservative)
    ref = _leaf_cones64(V, F, slots, owned, conservative)
    # leaves with a folded edge (two coincident faces of a soup) are noise in
    # FP32 and in the reference alike: only their radius is compared
    folded = np.array([any(e != -1 and _folded(V, slots[e]) for e in owned[i]) for i in range(len(F))])
    sentinel = (np.abs(ref[:, 3]) == np.pi) & ~folded
    np.testing.assert_array_equal(cones[sentinel, 3], np.float32(ref[sentinel, 3]))
    ok = ~sentinel & ~folded
    assert folded.sum() < len(F) // 10
    assert ok.sum() > len(F) // 10
    # half angle within 1e-5 rad, except for small angles where the FP32 dot
    # product's rounding near 1 (a few ulp, 6e-8 each) dominates: there the
    # cosines agree to 4 ulp; axis within 1e-5 rad; radius within 1e-6 relative
    dth = np.abs(cones[ok, 3] - ref[ok, 3])
    dcos = np.abs(np.cos(cones[ok, 3].astype(np.float64)) - np.cos(ref[ok, 3]))
    assert np.all((dth <= 1e-5) | (dcos <= 2.5e-7)), (dth.max(), dcos.max())
    a = cones[ok, :3].astype(np.float64)
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    ang = np.arctan2(np.linalg.norm(np.cross(a, ref[ok, :3]), axis=1), np.sum(a * ref[ok, :3], 1))
    assert np.max(ang) <= 1e-5
    np.testing.assert_allclose(cones[:, 4], ref[:, 4], rtol=1e-6, atol=1e-7)
    # enclosure: every owned edge's unit face normals lie within the cone (to
    # 1e-5 rad, or 4 ulp in cosine where the angle is small)
    for i in np.flatnonzero(ok):
        ax = cones[i, :3].astype(np.float64)
        ax /= np.linalg.norm(ax)
        for e in owned[i]:
            if e == -1:
                continue
            unit, _, _ = _edge_normals64(V, slots[e])
            for n in unit:
                c = np.clip(ax @ n, -1, 1)
                ang = np.arctan2(np.linalg.norm(np.cross(ax, n)), c)
                ha = float(cones[i, 3])
                assert ang <= ha + 1e-5 or c >= np.cos(ha) - 2.5e-7, (i, e, ang, ha)


# ------------------------------------------------------------ trig-free cone test
@pytest.fixture(scope="module")
def cone_check(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("conecheck") / "cone_fast_check.so")
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC",
                           "-o", out, os.path.join(ROOT, "tests", "cpp", "cone_fast_check.cpp")])
    L = ctypes.CDLL(out)
    L.cone_fast_check.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    return L


def _run_cone_check(L, oracle, cones, boxes, pts):
    ref = ctypes.cast(oracle.lib().orc_cone_overlap, ctypes.c_void_p)
    n = len(cones)
    cones = np.ascontiguousarray(cones, np.float32)
    boxes = np.ascontiguousarray(boxes, np.float32)
    pts = np.ascontiguousarray(pts, np.float32)
    res = []
    for mode in (0, 1, 2):
        m, x = ctypes.c_int64(), ctypes.c_int64()
        assert L.cone_fast_check(n, cones.ctypes.data, boxes.ctypes.data, pts.ctypes.data, mode, ref,
                                 ctypes.byref(m), ctypes.byref(x), None) == 0
        res.append((m.value, x.value))
    return res


def _unit(rng, n):
    d = rng.normal(size=(n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def test_cone_overlap_fast_keeps_every_reference_node(oracle, cone_check):
    """The trig-free test never drops a node the reference test keeps:
    random cones, boxes and view points, views nearly perpendicular to the
    axis (where |acos(C) - pi/2| is smallest), view points inside the box's
    bounding sphere, and the real nodes of a built tree."""
    rng = np.random.default_rng(11)
    n = 400000
    axis = _unit(rng, n)
    ha = rng.uniform(0.0, np.pi / 2, n) * (rng.random(n) < 0.8) + rng.uniform(0, 1e-3, n) * (rng.random(n) < 0.2)
    radius = np.exp(rng.uniform(np.log(1e-4), np.log(2.0), n))
    ctr = rng.uniform(-1, 1, (n, 3))
    ext = np.exp(rng.uniform(np.log(1e-3), np.log(0.5), (n, 3)))
    boxes = np.concatenate([ctr - ext, ctr + ext], 1)
    # view points: far, near, and nearly perpendicular to the axis through the centre
    dist = np.exp(rng.uniform(np.log(1e-3), np.log(5.0), n))
    dirs = _unit(rng, n)
    perp = np.cross(axis, _unit(rng, n))
    perp /= np.linalg.norm(perp, axis=1, keepdims=True)
    tilt = rng.normal(0, 1e-3, n)[:, None]
    view = np.where(rng.random(n)[:, None] < 0.5, dirs, perp * np.cos(tilt) + axis * np.sin(tilt))
    pts = ctr + view * dist[:, None]
    cones = np.concatenate([axis, ha[:, None], radius[:, None]], 1)
    for missed, excess in _run_cone_check(cone_check, oracle, cones, boxes, pts):
        assert missed == 0, missed
        assert excess < n // 100
    # half-angle tangency: views along the cone's edge (C = cos(pi/2 - ha - vha))
    m2 = 200000
    axis = _unit(rng, m2)
    ha = rng.uniform(0.0, np.pi / 2 - 1e-3, m2)
    radius = np.exp(rng.uniform(np.log(1e-3), np.log(0.3), m2))
    ctr = rng.uniform(-1, 1, (m2, 3))
    ext = np.full((m2, 3), 1e-3)
    l = np.exp(rng.uniform(np.log(0.5), np.log(5.0), m2))
    vha = np.arcsin(np.minimum(radius / l, 1.0))
    ang = np.pi / 2 - ha - vha + rng.normal(0, 1e-6, m2)
    side = np.cross(axis, _unit(rng, m2))
    side /= np.linalg.norm(side, axis=1, keepdims=True)
    v = axis * np.cos(ang)[:, None] + side * np.sin(ang)[:, None]
    pts = ctr - v * l[:, None]
    cones = np.concatenate([axis, ha[:, None], radius[:, None]], 1)
    for missed, _ in _run_cone_check(cone_check, oracle, cones, np.concatenate([ctr - ext, ctr + ext], 1), pts):
        assert missed == 0, missed
    # the nodes of a built tree (children's boxes and cones) against query points
    V, F = W.cut_torus(96, 48)
    t = oracle.build_tree(V, F)
    NN = len(t.nodes)
    pick = rng.integers(1, NN, 300000)
    lo, hi = W.padded_box(V, 0.25)
    pts = W.uniform_points(len(pick), lo, hi, 12)
    for missed, excess in _run_cone_check(cone_check, oracle, t.cones[pick], t.boxes[pick], pts):
        assert missed == 0, missed


# ------------------------------------------------------------ 2D sampler
def _path_probabilities2d(t, c, r):
    """Exact descent probabilities of sample.cuh:23-92 for scene<2> (Green
    weight |log(max(r, 1e-2))| / 2 pi, scene.cuh:605-611), in float64."""
    N = t.n_obj

    def inter(box):
        cl = np.maximum(box[:2], np.minimum(c, box[3:5]))
        return np.sum((cl - c) ** 2) <= r * r

    def weight(box):
        cen = 0.5 * (box[:2].astype(np.float64) + box[3:5])
        return abs(np.log(max(np.linalg.norm(c - cen), 1e-2)) / (2 * np.pi))

    probs = {}
    stack = [(0, 1.0)]
    while stack:
        n, p = stack.pop()
        if n >= N - 1 or N == 1:
            probs[int(t.nodes[n, 3])] = probs.get(int(t.nodes[n, 3]), 0.0) + p
            continue
        l, rr = int(t.nodes[n, 1]), int(t.nodes[n, 2])
        wl = weight(t.boxes[l]) if inter(t.boxes[l]) else 0.0
        wr = weight(t.boxes[rr]) if inter(t.boxes[rr]) else 0.0
        if wl + wr > 0:
            if wl > 0:
                stack.append((l, p * wl / (wl + wr)))
            if wr > 0:
                stack.append((rr, p * wr / (wl + wr)))
    return probs


def test_sample2d_distribution(oracle):
    V, S = W.circle_in_square(96, 24, sigma=0.01)
    t = oracle.build_tree2d(V, S)
    c = np.float32([0.45, 0.1])
    r = np.float32(0.35)
    n = 200000
    u = np.random.default_rng(5).random((n, 3)).astype(np.float32)
    idx, pdf, pt = oracle.sample2d(t, np.tile(c, (n, 1)), np.full(n, r), u)
    probs = _path_probabilities2d(t, c.astype(np.float64), float(r))
    length = np.linalg.norm(V[S[:, 1]].astype(np.float64) - V[S[:, 0]], axis=1)
    ok = idx >= 0
    assert ok.mean() > 0.5
    exp_pdf = np.array([probs[i] for i in idx[ok]]) / length[idx[ok]]
    np.testing.assert_allclose(pdf[ok], exp_pdf, rtol=2e-4)
    assert np.all(pdf[~ok] == 0)
    keys = sorted(probs)
    cnt = np.array([np.sum(idx == k) for k in keys], np.float64)
    p = np.array([probs[k] for k in keys])
    exp = n * p
    fails = np.array([cnt[i] == 0 and exp[i] > 50 for i in range(len(keys))])
    miss = n - ok.sum()
    exp_miss = exp[fails].sum() + n * (1.0 - p.sum())
    chi = np.sum((cnt[~fails] - exp[~fails]) ** 2 / exp[~fails]) + (miss - exp_miss) ** 2 / max(exp_miss, 1)
    assert chi < len(keys) + 6 * np.sqrt(2 * len(keys)), chi
    # sample points lie on their segments, at parameter u[1]
    seg = V[S[idx[ok]]].astype(np.float64)
    want = seg[:, 0] + u[ok, 1:2].astype(np.float64) * (seg[:, 1] - seg[:, 0])
    assert np.max(np.abs(pt[ok] - want)) < 1e-5


# ------------------------------------------------------------ 2D ray
def _bf_ray2d(V, S, o, d, s_lo, s_hi):
    """FP64 brute force of scene.cuh:545-576 (segment parameter s accepted in
    [s_lo, s_hi], t >= 0): nearest t per ray, +inf when none."""
    V = V.astype(np.float64)
    p0, p1 = V[S[:, 0]], V[S[:, 1]]
    sd = p1 - p0
    o = o.astype(np.float64)[:, None, :]
    d = d.astype(np.float64)[:, None, :]
    D = d[..., 0] * (-sd[None, :, 1]) + d[..., 1] * sd[None, :, 0]
    w = p0[None] - o
    with np.errstate(divide="ignore", invalid="ignore"):
        t = (w[..., 0] * (-sd[None, :, 1]) - w[..., 1] * (-sd[None, :, 0])) / D
        s = (d[..., 0] * w[..., 1] - d[..., 1] * w[..., 0]) / D
    ok = (np.abs(D) >= 1.1920928955078125e-07) & (s >= s_lo) & (s <= s_hi) & (t >= 0)
    tt = np.where(ok, t, np.inf)
    j = tt.argmin(1)
    rows = np.arange(len(j))
    # sine of the ray / segment angle at the nearest hit: FP32 rounding of t
    # grows as 1 / sin (a grazing ray)
    sin = np.abs(D[rows, j]) / (np.linalg.norm(d[:, 0], axis=1) * np.linalg.norm(sd[j], axis=1))
    return tt[rows, j], sin


@pytest.mark.parametrize("which", ["cis", "soup"])
def test_ray2d_vs_fp64_bruteforce(oracle, which):
    V, S = W.circle_in_square(200, 30, sigma=0.02) if which == "cis" else W.polyline_soup(300, 500)
    t = oracle.build_tree2d(V, S)
    rng = np.random.default_rng(4)
    n = 3000
    o = rng.uniform(-1.2, 1.2, (n, 2)).astype(np.float32)
    ang = rng.uniform(0, 2 * np.pi, n)
    d = np.stack([np.cos(ang), np.sin(ang)], 1).astype(np.float32)
    hit, tt, _, _ = oracle.ray2d(t, o, d)
    strict, sin = _bf_ray2d(V, S, o, d, 0.0, 1.0)            # hits inside a segment
    loose, _ = _bf_ray2d(V, S, o, d, -1e-3, 1.0 + 1e-3)      # the reference's tolerance
    # FP32 t: a few ulp of the coordinates, divided by the sine of the angle
    with np.errstate(invalid="ignore"):
        tol = 1e-6 * (1.0 + np.abs(strict)) / np.maximum(sin, 1e-3)
    # a hit inside a segment is always found, and nothing nearer than the
    # tolerant brute force's nearest is reported
    fin = np.isfinite(strict)
    assert np.all(hit[fin])
    assert np.all(tt[fin] <= strict[fin] + tol[fin])
    assert np.all(np.isfinite(loose[hit]))
    assert np.all(tt[hit] >= loose[hit] - 1e-3 * (1.0 + np.abs(loose[hit])))
    # where the two brute forces agree, the hit distance matches
    same = fin & (np.abs(strict - np.where(np.isfinite(loose), loose, 0.0)) <= tol)
    assert np.all(np.abs(tt[same] - strict[same]) <= tol[same])
    assert same.mean() > 0.5


# ------------------------------------------------------------ walk on stars
def test_wost_mixed_slab_kat(oracle):
    """Dirichlet g = x y on the four sides of the slab [-1,1]^2 x [-h,h] and
    zero Neumann flux on its caps: u = x y exactly, so the estimator's mean
    at each start is x0.x * x0.y.  Exercises reflection off the Neumann caps
    (ray hits) and the silhouette star radius."""
    V, Fs, Fc = W.slab_box(12, 0.25)
    D = oracle.build_tree(V, Fs)
    N = oracle.build_tree(V, Fc)
    starts = np.float32([[0.3, -0.5, 0.1], [-0.6, 0.4, -0.2], [0.05, 0.7, 0.22]])
    k = 6000
    x0 = np.repeat(starts, k, axis=0)
    vals, steps = oracle.wost(D, N, x0, max_steps=256, eps=1e-3, seed=5)
    assert np.all(steps < 256)
    v = vals.reshape(3, k)
    for s, row in zip(starts, v):
        exact = float(s[0]) * float(s[1])
        se = row.std() / np.sqrt(k)
        assert abs(row.mean() - exact) < 4 * se + 2e-3, (s, row.mean(), exact, se)
    # the caps are reached: some walks take many more steps than pure-Dirichlet ones
    assert np.median(steps) > 3
