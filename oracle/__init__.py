"""ORACLE -- test infrastructure only.

ctypes bindings to ``liboracle.so``, the plain-C restatement of the reference
SNCH-LBVH algorithm (``oracle/snch_oracle.c``).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module, and only as the checker; the library's product path never
touches it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

U32 = np.uint32
I32 = np.int32
F32 = np.float32


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _setup(_lib)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class OrcTree(ctypes.Structure):
    _fields_ = [
        ("n_obj", ctypes.c_uint32),
        ("nodes", ctypes.c_void_p),
        ("boxes", ctypes.c_void_p),
        ("cones", ctypes.c_void_p),
        ("V", ctypes.c_void_p),
        ("F", ctypes.c_void_p),
        ("slots", ctypes.c_void_p),
        ("owned", ctypes.c_void_p),
    ]


def _setup(L):
    vp, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
    L.orc_edge_table.argtypes = [vp, i32, vp, vp, vp]
    L.orc_edge_table.restype = i32
    L.orc_build.argtypes = [vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int]
    L.orc_build.restype = ctypes.c_int
    L.orc_leaf_box_cone.argtypes = [vp, vp, i32, vp, vp, vp, vp, ctypes.c_int]
    L.orc_morton3.argtypes = [f32, f32, f32]
    L.orc_morton3.restype = ctypes.c_uint32
    L.orc_expand_bits.argtypes = [ctypes.c_uint32]
    L.orc_expand_bits.restype = ctypes.c_uint32
    L.orc_cone_merge.argtypes = [vp] * 6 + [ctypes.c_int]
    L.orc_cone_overlap.argtypes = [vp, vp, vp, f32]
    L.orc_cone_overlap.restype = ctypes.c_int
    L.orc_math.argtypes = [ctypes.c_int, vp, vp, vp, i32]
    L.orc_closest_point.argtypes = [vp, vp, i64, vp, vp, vp, ctypes.c_int]
    L.orc_closest_silhouette.argtypes = [vp, vp, i64, ctypes.c_int, vp, vp, vp, ctypes.c_int]
    L.orc_ray.argtypes = [vp, vp, vp, vp, i64, ctypes.c_int, vp, vp, vp, vp, vp, ctypes.c_int]
    L.orc_sample_in_sphere.argtypes = [vp, vp, vp, vp, i64, vp, vp, vp, ctypes.c_int]
    L.orc_bf_closest_point.argtypes = [vp, vp, i32, vp, i64, vp, vp, ctypes.c_int]
    L.orc_bf_closest_silhouette.argtypes = [vp, vp, i32, vp, i64, ctypes.c_int, vp, vp, ctypes.c_int]
    L.orc_bf_ray.argtypes = [vp, vp, i32, vp, vp, vp, i64, vp, vp, ctypes.c_int]
    L.orc_tri_dist64.argtypes = [vp, vp, vp, vp, i64, vp]
    L.orc_tri_dist32.argtypes = [vp, vp, vp, vp, i64, vp]
    L.orc_edge_sil32.argtypes = [vp, vp, vp, vp, i64, ctypes.c_int, vp]


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def edge_table(F):
    """Reference edge ids per (triangle, local edge), edge slots (opp0, lo, hi,
    opp1) and per-triangle owned edges (scene.cuh:1135-1225)."""
    F = _c(F, I32).reshape(-1, 3)
    nF = F.shape[0]
    tri_edge = np.empty((nF, 3), I32)
    slots = np.empty((max(3 * nF, 1), 4), I32)
    owned = np.empty((nF, 3), I32)
    E = lib().orc_edge_table(_p(F), nF, _p(tri_edge), _p(slots), _p(owned))
    return tri_edge, slots[:E].copy(), owned


class Tree:
    """A tree in the reference layout (bvh.cuh:27-33, 380-613)."""

    def __init__(self, V, F, slots, owned, nodes, boxes, cones, sorted_codes=None, sorted_idx=None,
                 collision=False, scene_box=None):
        self.V, self.F, self.slots, self.owned = V, F, slots, owned
        self.nodes, self.boxes, self.cones = nodes, boxes, cones
        self.sorted_codes, self.sorted_idx = sorted_codes, sorted_idx
        self.collision = collision
        self.scene_box = scene_box
        self.n_obj = F.shape[0]
        self._s = OrcTree(self.n_obj, _p(nodes), _p(boxes), _p(cones), _p(V), _p(F), _p(slots), _p(owned))

    @property
    def ref(self):
        return ctypes.byref(self._s)


CONSERVATIVE_RADIUS = 1


def build_tree(V, F, edges=None, conservative_radius=True):
    """Full reference construct() on the CPU.  conservative_radius selects the
    enclosing cone radius (default) instead of the reference's midpoint radius."""
    V = _c(V, F32).reshape(-1, 3)
    F = _c(F, I32).reshape(-1, 3)
    if edges is None:
        edges = edge_table(F)
    _, slots, owned = edges
    slots = _c(slots, I32)
    owned = _c(owned, I32)
    N = F.shape[0]
    NN = 2 * N - 1
    nodes = np.empty((NN, 4), U32)
    boxes = np.empty((NN, 6), F32)
    cones = np.empty((NN, 5), F32)
    codes = np.empty(N, U32)
    sidx = np.empty(N, U32)
    sbox = np.empty(6, F32)
    col = lib().orc_build(_p(V), _p(F), N, _p(slots), _p(owned), _p(nodes), _p(boxes), _p(cones), _p(codes),
                          _p(sidx), _p(sbox), CONSERVATIVE_RADIUS if conservative_radius else 0)
    if col < 0:
        raise ValueError("empty mesh")
    return Tree(V, F, slots, owned, nodes, boxes, cones, codes, sidx, bool(col), sbox)


def leaf_box_cone(V, F, edges=None, conservative_radius=True):
    V = _c(V, F32).reshape(-1, 3)
    F = _c(F, I32).reshape(-1, 3)
    if edges is None:
        edges = edge_table(F)
    _, slots, owned = edges
    n = F.shape[0]
    boxes = np.empty((n, 6), F32)
    cones = np.empty((n, 5), F32)
    lib().orc_leaf_box_cone(_p(V), _p(F), n, _p(_c(slots, I32)), _p(_c(owned, I32)), _p(boxes), _p(cones),
                            1 if conservative_radius else 0)
    return boxes, cones


def math(fn, x, y=None):
    """fn: 0 acos, 1 asin, 2 atan2(x, y), 3 sin, 4 cos (deterministic FP32 set)."""
    x = _c(x, F32).ravel()
    y = _c(y if y is not None else np.zeros_like(x), F32).ravel()
    out = np.empty_like(x)
    lib().orc_math(fn, _p(x), _p(y), _p(out), x.size)
    return out


def cone_merge(a, b, oa, ob, on, conservative_radius=True):
    out = np.empty(5, F32)
    args = [_c(v, F32) for v in (a, b, oa, ob, on)]
    lib().orc_cone_merge(*[_p(v) for v in args], _p(out), 1 if conservative_radius else 0)
    return out


def cone_overlap(cone, o, box, dist2):
    return bool(lib().orc_cone_overlap(_p(_c(cone, F32)), _p(_c(o, F32)), _p(_c(box, F32)), float(dist2)))


def closest_point(tree, q, nthreads=0, counts=False):
    q = _c(q, F32).reshape(-1, 3)
    n = q.shape[0]
    idx = np.empty(n, U32)
    dist = np.empty(n, F32)
    cnt = np.zeros(3, np.uint64) if counts else None
    lib().orc_closest_point(tree.ref, _p(q), n, _p(idx), _p(dist), _p(cnt), nthreads)
    return (idx, dist, cnt) if counts else (idx, dist)


def closest_silhouette(tree, q, flip=False, nthreads=0, counts=False):
    q = _c(q, F32).reshape(-1, 3)
    n = q.shape[0]
    edge = np.empty(n, U32)
    dist = np.empty(n, F32)
    cnt = np.zeros(3, np.uint64) if counts else None
    lib().orc_closest_silhouette(tree.ref, _p(q), n, int(bool(flip)), _p(edge), _p(dist), _p(cnt), nthreads)
    return (edge, dist, cnt) if counts else (edge, dist)


def wost(dirichlet, neumann, x0, max_steps=64, eps=1e-3, seed=0, nthreads=0):
    """Walk on stars, one walk at a time (oracle/snch_oracle.c orc_wost): the
    library's per-walk arithmetic and random streams over the oracle's
    queries.  dirichlet / neumann: Tree (neumann may be None)."""
    L = lib()
    if not getattr(L, "_wost_declared", False):
        L.orc_wost.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint32,
                               ctypes.c_float, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L._wost_declared = True
    x0 = _c(x0, F32).reshape(-1, 3)
    n = x0.shape[0]
    out = np.empty(n, F32)
    steps = np.empty(n, U32)
    L.orc_wost(dirichlet.ref, neumann.ref if neumann is not None else None, _p(x0), n, max_steps, float(eps),
               int(seed), _p(out), _p(steps), nthreads)
    return out, steps


def ray(tree, o, d, tmax=None, any_hit=False, nthreads=0, counts=False):
    o = _c(o, F32).reshape(-1, 3)
    d = _c(d, F32).reshape(-1, 3)
    n = o.shape[0]
    tmax = _c(np.full(n, np.inf) if tmax is None else tmax, F32)
    idx = np.empty(n, U32)
    t = np.empty(n, F32)
    uv = np.empty((n, 2), F32)
    hit = np.empty(n, np.uint8)
    cnt = np.zeros(3, np.uint64) if counts else None
    lib().orc_ray(tree.ref, _p(o), _p(d), _p(tmax), n, int(bool(any_hit)), _p(idx), _p(t), _p(uv), _p(hit),
                  _p(cnt), nthreads)
    out = (hit.astype(bool), t, uv, idx)
    return out + (cnt,) if counts else out


def sample_in_sphere(tree, c, r, u, nthreads=0):
    c = _c(c, F32).reshape(-1, 3)
    n = c.shape[0]
    r = _c(r, F32).ravel()
    u = _c(u, F32).reshape(n, 3)
    idx = np.empty(n, I32)
    pdf = np.empty(n, F32)
    pt = np.empty((n, 3), F32)
    lib().orc_sample_in_sphere(tree.ref, _p(c), _p(r), _p(u), n, _p(idx), _p(pdf), _p(pt), nthreads)
    return idx, pdf, pt


def bf_closest_point(V, F, q, nthreads=0):
    V = _c(V, F32).reshape(-1, 3)
    F = _c(F, I32).reshape(-1, 3)
    q = _c(q, F32).reshape(-1, 3)
    n = q.shape[0]
    idx = np.empty(n, U32)
    dist = np.empty(n, np.float64)
    lib().orc_bf_closest_point(_p(V), _p(F), F.shape[0], _p(q), n, _p(idx), _p(dist), nthreads)
    return idx, dist


def bf_closest_silhouette(V, slots, q, flip=False, nthreads=0):
    V = _c(V, F32).reshape(-1, 3)
    slots = _c(slots, I32).reshape(-1, 4)
    q = _c(q, F32).reshape(-1, 3)
    n = q.shape[0]
    edge = np.empty(n, U32)
    dist = np.empty(n, np.float64)
    lib().orc_bf_closest_silhouette(_p(V), _p(slots), slots.shape[0], _p(q), n, int(bool(flip)), _p(edge),
                                    _p(dist), nthreads)
    return edge, dist


def bf_ray(V, F, o, d, tmax=None, nthreads=0):
    V = _c(V, F32).reshape(-1, 3)
    F = _c(F, I32).reshape(-1, 3)
    o = _c(o, F32).reshape(-1, 3)
    d = _c(d, F32).reshape(-1, 3)
    n = o.shape[0]
    tmax = _c(np.full(n, np.inf) if tmax is None else tmax, F32)
    idx = np.empty(n, U32)
    t = np.empty(n, np.float64)
    lib().orc_bf_ray(_p(V), _p(F), F.shape[0], _p(o), _p(d), _p(tmax), n, _p(idx), _p(t), nthreads)
    return idx, t


def tri_dist(V, F, tri, q, f64=False):
    V = _c(V, F32).reshape(-1, 3)
    F = _c(F, I32).reshape(-1, 3)
    tri = _c(tri, U32).ravel()
    q = _c(q, F32).reshape(-1, 3)
    out = np.empty(tri.size, np.float64 if f64 else F32)
    (lib().orc_tri_dist64 if f64 else lib().orc_tri_dist32)(_p(V), _p(F), _p(tri), _p(q), tri.size, _p(out))
    return out


def edge_sil_dist(V, slots, edge, q, flip=False):
    """FP32 reference silhouette distance of the given edge (inf if not a silhouette)."""
    V = _c(V, F32).reshape(-1, 3)
    slots = _c(slots, I32).reshape(-1, 4)
    edge = _c(edge, U32).ravel()
    q = _c(q, F32).reshape(-1, 3)
    out = np.empty(edge.size, F32)
    lib().orc_edge_sil32(_p(V), _p(slots), _p(edge), _p(q), edge.size, int(bool(flip)), _p(out))
    return out


def edge_sil_dist64(V, slots, edge, q, flip=False):
    """FP64 silhouette distance of the given edge under the reference predicate (inf if not one)."""
    V = _c(V, F32).reshape(-1, 3)
    slots = _c(slots, I32).reshape(-1, 4)
    edge = _c(edge, U32).ravel()
    q = _c(q, F32).reshape(-1, 3)
    out = np.empty(edge.size, np.float64)
    L = lib()
    L.orc_edge_sil64.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
    L.orc_edge_sil64(_p(V), _p(slots), _p(edge), _p(q), edge.size, int(bool(flip)), _p(out))
    return out


# ---------------------------------------------------------------- 2D (scene<2>)
class OrcTree2D(ctypes.Structure):
    _fields_ = OrcTree._fields_


def silhouettes2d(S, nV):
    """Silhouette-vertex slots (prev, v, next, -1) per vertex and the 2 owned
    vertices per segment (scene.cuh:631-672)."""
    S = _c(S, I32).reshape(-1, 2)
    slots = np.empty((nV, 4), I32)
    owned = np.empty((S.shape[0], 2), I32)
    L = lib()
    L.orc_silhouettes2d.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
    L.orc_silhouettes2d(_p(S), S.shape[0], nV, _p(slots), _p(owned))
    return slots, owned


class Tree2D(Tree):
    def __init__(self, V, S, slots, owned, nodes, boxes, cones, codes, sidx, collision):
        self.V, self.F, self.slots, self.owned = V, S, slots, owned
        self.nodes, self.boxes, self.cones = nodes, boxes, cones
        self.sorted_codes, self.sorted_idx, self.collision = codes, sidx, collision
        self.n_obj = S.shape[0]
        self._s = OrcTree2D(self.n_obj, _p(nodes), _p(boxes), _p(cones), _p(V), _p(S), _p(slots), _p(owned))


def build_tree2d(V, S, conservative_radius=True):
    V = _c(V, F32).reshape(-1, 2)
    S = _c(S, I32).reshape(-1, 2)
    slots, owned = silhouettes2d(S, V.shape[0])
    N = S.shape[0]
    NN = 2 * N - 1
    nodes = np.empty((NN, 4), U32)
    boxes = np.empty((NN, 6), F32)
    cones = np.empty((NN, 5), F32)
    codes = np.empty(N, U32)
    sidx = np.empty(N, U32)
    L = lib()
    L.orc_build2d.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_int32] + [ctypes.c_void_p] * 7 + [ctypes.c_int]
    col = L.orc_build2d(_p(V), _p(S), N, _p(slots), _p(owned), _p(nodes), _p(boxes), _p(cones), _p(codes), _p(sidx),
                        1 if conservative_radius else 0)
    if col < 0:
        raise ValueError("empty mesh")
    return Tree2D(V, S, slots, owned, nodes, boxes, cones, codes, sidx, bool(col))


def closest_point2d(tree, q):
    q = _c(q, F32).reshape(-1, 2)
    n = q.shape[0]
    idx = np.empty(n, U32)
    dist = np.empty(n, F32)
    L = lib()
    L.orc_closest_point2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    L.orc_closest_point2d(tree.ref, _p(q), n, _p(idx), _p(dist))
    return idx, dist


def closest_silhouette2d(tree, q, flip=False):
    q = _c(q, F32).reshape(-1, 2)
    n = q.shape[0]
    sil = np.empty(n, U32)
    dist = np.empty(n, F32)
    L = lib()
    L.orc_closest_silhouette2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                           ctypes.c_void_p, ctypes.c_void_p]
    L.orc_closest_silhouette2d(tree.ref, _p(q), n, int(bool(flip)), _p(sil), _p(dist))
    return sil, dist


def ray2d(tree, o, d, tmax=None):
    o = _c(o, F32).reshape(-1, 2)
    d = _c(d, F32).reshape(-1, 2)
    n = o.shape[0]
    tmax = _c(np.full(n, np.inf) if tmax is None else tmax, F32)
    idx = np.empty(n, U32)
    t = np.empty(n, F32)
    s = np.empty(n, F32)
    hit = np.empty(n, np.uint8)
    L = lib()
    L.orc_ray2d.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64] + [ctypes.c_void_p] * 4
    L.orc_ray2d(tree.ref, _p(o), _p(d), _p(tmax), n, _p(idx), _p(t), _p(s), _p(hit))
    return hit.astype(bool), t, s, idx


def sample2d(tree, c, r, u):
    c = _c(c, F32).reshape(-1, 2)
    n = c.shape[0]
    r = _c(r, F32).ravel()
    u = _c(u, F32).reshape(n, 3)
    idx = np.empty(n, I32)
    pdf = np.empty(n, F32)
    pt = np.empty((n, 2), F32)
    L = lib()
    L.orc_sample2d.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64] + [ctypes.c_void_p] * 3
    L.orc_sample2d(tree.ref, _p(c), _p(r), _p(u), n, _p(idx), _p(pdf), _p(pt))
    return idx, pdf, pt


def bf_closest_point2d(V, S, q):
    V = _c(V, F32).reshape(-1, 2)
    S = _c(S, I32).reshape(-1, 2)
    q = _c(q, F32).reshape(-1, 2)
    n = q.shape[0]
    idx = np.empty(n, U32)
    dist = np.empty(n, np.float64)
    L = lib()
    L.orc_bf_closest_point2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    L.orc_bf_closest_point2d(_p(V), _p(S), S.shape[0], _p(q), n, _p(idx), _p(dist))
    return idx, dist


def bf_closest_silhouette2d(V, slots, q, flip=False):
    V = _c(V, F32).reshape(-1, 2)
    slots = _c(slots, I32).reshape(-1, 4)
    q = _c(q, F32).reshape(-1, 2)
    n = q.shape[0]
    sil = np.empty(n, U32)
    dist = np.empty(n, np.float64)
    L = lib()
    L.orc_bf_closest_silhouette2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                              ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    L.orc_bf_closest_silhouette2d(_p(V), _p(slots), slots.shape[0], _p(q), n, int(bool(flip)), _p(sil), _p(dist))
    return sil, dist


# ---------------------------------------------------------------- generic queries
def _generic(tree, q6, kind, max_out):
    dim = 2 if isinstance(tree, Tree2D) else 3
    q6 = _c(q6, F32).reshape(-1, 6)
    n = q6.shape[0]
    counts = np.empty(n, U32)
    idx = np.empty((n, max(max_out, 1)), U32)
    t = np.empty((n, max(max_out, 1)), F32) if kind == 1 else None
    L = lib()
    L.orc_generic_query.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    L.orc_generic_query(tree.ref, dim, kind, _p(q6), n, max(max_out, 1), _p(counts), _p(idx), _p(t), 0)
    return counts, idx[:, :max_out], (t[:, :max_out] if t is not None else None)


def _pad6(a, b, dim):
    a = np.asarray(a, np.float32).reshape(-1, dim)
    b = np.asarray(b, np.float32).reshape(-1, dim)
    z = np.zeros((a.shape[0], 3 - dim), np.float32)
    return np.concatenate([a, z, b, z], 1)


def overlap(tree, lo, hi, max_out):
    """query_device(bvh, overlaps(box), out, max_out) (query.cuh:176-231) per
    query box [lo, hi]: (total found, first max_out ids in the reference's
    visit order)."""
    dim = 2 if isinstance(tree, Tree2D) else 3
    counts, idx, _ = _generic(tree, _pad6(lo, hi, dim), 0, max_out)
    return counts, idx


def line_intersect(tree, o, d, max_out):
    """query_device(bvh, line_intersect(line(o, d)), test, out, max_out)
    (query.cuh:12-77), the element test being the scene's ray test without
    t >= 0: (total found, first max_out ids, their line parameters t)."""
    dim = 2 if isinstance(tree, Tree2D) else 3
    return _generic(tree, _pad6(o, d, dim), 1, max_out)
