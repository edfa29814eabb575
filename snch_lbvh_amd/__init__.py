"""SNCH-LBVH for MI355X: Python mirror of the reference's ``lbvh::scene`` API.

The compute path is ``libsnch.so`` (hand-written HIP kernels for gfx950 behind
the C ABI in ``include/snch.h``); this module only binds it with ctypes and
passes device pointers of torch CUDA tensors.  There is no CPU fallback: if
the library or a GPU is missing, construction raises.

Reference API -> this module
  lbvh::scene<3>(vb, ve, ib, ie); compute_silhouettes()  -> Scene(V, F)
  scene.build_bvh()                                    -> Scene.build_bvh()
  query_device(bvh, nearest(p), distance_calculator)   -> Scene.closest_point(q)
  query_device(bvh, nearest_silhouette(p, flip), ...)  -> Scene.closest_silhouette(q, flip)
  query_device(bvh, ray_intersect<TestOnly>(r, tmax))  -> Scene.ray_intersect(o, d, tmax, any_hit)
  sample_object_in_sphere(...) + sample_on_object(...) -> Scene.sample_in_sphere(c, r, u)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsnch.so")
_lib = None

SNCH_OK = 0
ERRORS = {1: "invalid argument", 2: "empty mesh", 3: "HIP error", 4: "not built", 5: "stack overflow"}
BUILD_REFERENCE_RADIUS = 1
BUILD_TIME_PHASES = 2
BUILD_TOPS_IN_MEMORY = 4  # layouts of the refit above 16384-leaf subtrees, forced for tests
BUILD_TOPS_CLIMB = 8
_TOP_LAYOUT = {"auto": 0, "memory": BUILD_TOPS_IN_MEMORY, "climb": BUILD_TOPS_CLIMB}


class SnchError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"snch error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


def lib():
    """Load libsnch.so (import torch first so that one HIP runtime serves both)."""
    global _lib
    if _lib is None:
        import torch  # noqa: F401  (shares its libamdhip64 with libsnch)

        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m snch_lbvh_amd.build` first")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


class BuildOpts(ctypes.Structure):
    _fields_ = [("flags", ctypes.c_uint32), ("leaf_size", ctypes.c_uint32)]


class QueryOpts(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_uint32), ("coherent", ctypes.c_uint32), ("stats", ctypes.c_void_p),
                ("order", ctypes.c_void_p), ("stack_depth", ctypes.c_uint32), ("queries_per_wave", ctypes.c_uint32),
                ("seed_is_cap", ctypes.c_uint32), ("seed_retry", ctypes.c_uint32)]


class DeviceView(ctypes.Structure):
    """snch_device_view: a built mesh's tree for users' own HIP kernels
    (include/snch_lbvh/device.hpp)."""
    _fields_ = [("dim", ctypes.c_uint32), ("n_leaves", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("qbase", ctypes.c_uint32), ("tnode", ctypes.c_void_p), ("tcone", ctypes.c_void_p),
                ("tleaf", ctypes.c_void_p), ("tedge", ctypes.c_void_p), ("nbox", ctypes.c_void_p),
                ("ncone", ctypes.c_void_p), ("V", ctypes.c_void_p), ("F", ctypes.c_void_p)]


class TuneResult(ctypes.Structure):
    _fields_ = [("leaf_size", ctypes.c_uint32), ("block_size", ctypes.c_uint32), ("stack_depth", ctypes.c_uint32),
                ("queries_per_wave", ctypes.c_uint32), ("ms", ctypes.c_float), ("ms_default", ctypes.c_float),
                ("n_tried", ctypes.c_uint32)]


EXPORTED = [
    "snch_last_error", "snch_version", "snch_context_create", "snch_context_destroy", "snch_scene_create_tri",
    "snch_scene_create_tri_batch", "snch_scene_create_seg", "snch_scene_destroy", "snch_scene_mesh_count",
    "snch_scene_info", "snch_scene_build", "snch_scene_build_time", "snch_scene_build_phases",
    "snch_batch_closest_point",
    "snch_batch_closest_silhouette", "snch_batch_autotune", "snch_scene_export_mesh", "snch_scene_device_view",
    "snch_query_order", "snch_closest_point", "snch_closest_point_ex", "snch_closest_silhouette",
    "snch_closest_silhouette_ex", "snch_closest_silhouette_noalloc", "snch_ray", "snch_ray_ex",
    "snch_sample_in_sphere", "snch_sample_on_object", "snch_overlap", "snch_line_intersect", "snch_scene_export", "snch_scene_silhouettes",
    "snch_device_malloc", "snch_device_free", "snch_memcpy_h2d", "snch_memcpy_d2h", "snch_stream_synchronize",
    "snch_mlp_num_params", "snch_mlp_create", "snch_mlp_destroy", "snch_mlp_params", "snch_mlp_train",
    "snch_mlp_set_loss",
    "snch_mlp_gradient", "snch_mlp_predict", "snch_wost_create", "snch_wost_destroy", "snch_wost_solve",
]


def _declare(L):
    vp, u32, i32, c_int = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int
    L.snch_last_error.restype = ctypes.c_char_p
    L.snch_context_create.argtypes = [c_int, ctypes.POINTER(vp)]
    L.snch_context_destroy.argtypes = [vp]
    L.snch_scene_create_tri.argtypes = [vp, vp, u32, vp, u32, ctypes.POINTER(vp)]
    L.snch_scene_create_seg.argtypes = [vp, vp, u32, vp, u32, ctypes.POINTER(vp)]
    L.snch_scene_create_tri_batch.argtypes = [vp, u32, vp, vp, vp, vp, ctypes.POINTER(vp)]
    L.snch_scene_mesh_count.argtypes = [vp, vp]
    L.snch_batch_closest_point.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.snch_batch_closest_silhouette.argtypes = [vp, vp, vp, c_int, vp, vp, vp, vp, vp]
    L.snch_batch_autotune.argtypes = [vp, u32, vp, vp, u32, c_int, vp, vp]
    L.snch_scene_export_mesh.argtypes = [vp, u32, vp, vp, vp, vp, vp]
    L.snch_scene_device_view.argtypes = [vp, u32, vp]
    L.snch_scene_destroy.argtypes = [vp]
    L.snch_scene_info.argtypes = [vp, vp, vp, vp]
    L.snch_scene_build.argtypes = [vp, vp, vp]
    L.snch_scene_build_time.argtypes = [vp, vp]
    L.snch_scene_build_phases.argtypes = [vp, vp]
    L.snch_query_order.argtypes = [vp, vp, u32, vp, vp]
    L.snch_closest_point.argtypes = [vp, vp, u32, vp, vp, vp]
    L.snch_closest_silhouette.argtypes = [vp, vp, u32, c_int, vp, vp, vp, vp]
    L.snch_closest_silhouette_ex.argtypes = [vp, vp, u32, c_int, vp, vp, vp, vp, vp]
    L.snch_closest_silhouette_noalloc.argtypes = [vp, vp, u32, c_int, vp, vp, vp, vp]
    L.snch_closest_point_ex.argtypes = [vp, vp, u32, vp, vp, vp, vp]
    L.snch_ray_ex.argtypes = [vp, vp, vp, vp, u32, c_int, vp, vp, vp, vp, vp, vp]
    L.snch_ray.argtypes = [vp, vp, vp, vp, u32, c_int, vp, vp, vp, vp, vp]
    L.snch_sample_in_sphere.argtypes = [vp, vp, vp, vp, u32, vp, vp, vp, vp]
    L.snch_sample_on_object.argtypes = [vp, vp, vp, u32, vp, vp]
    L.snch_overlap.argtypes = [vp, vp, vp, u32, u32, vp, vp, vp]
    L.snch_line_intersect.argtypes = [vp, vp, vp, u32, u32, vp, vp, vp, vp]
    L.snch_scene_export.argtypes = [vp, vp, vp, vp, vp, vp]
    L.snch_scene_silhouettes.argtypes = [vp, vp, vp]
    L.snch_mlp_create.argtypes = [vp, vp, vp, ctypes.c_float, ctypes.c_uint64, ctypes.POINTER(vp)]
    L.snch_mlp_destroy.argtypes = [vp]
    L.snch_mlp_params.argtypes = [vp, vp, vp, vp, vp]
    L.snch_mlp_train.argtypes = [vp, vp, vp, u32, u32, u32, ctypes.c_float, vp, vp]
    L.snch_mlp_gradient.argtypes = [vp, vp, vp, u32, vp, vp, vp]
    L.snch_mlp_set_loss.argtypes = [vp, ctypes.c_float]
    L.snch_mlp_predict.argtypes = [vp, vp, u32, c_int, ctypes.c_float, ctypes.c_float, vp, vp]
    L.snch_wost_create.argtypes = [vp, vp, vp, u32, ctypes.POINTER(vp)]
    L.snch_wost_destroy.argtypes = [vp]
    L.snch_wost_solve.argtypes = [vp, vp, u32, u32, ctypes.c_float, ctypes.c_uint64, c_int, vp, vp, vp]
    for name in EXPORTED:
        if name != "snch_last_error":
            getattr(L, name).restype = i32
    L.snch_mlp_num_params.restype = u32


def check(rc):
    if rc != SNCH_OK:
        raise SnchError(rc, lib().snch_last_error().decode())


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


_contexts = {}


def context(device=0):
    if device not in _contexts:
        c = ctypes.c_void_p()
        check(lib().snch_context_create(device, ctypes.byref(c)))
        _contexts[device] = c
    return _contexts[device]


def _dev_f32(x, cols, device):
    import torch

    t = torch.as_tensor(x, dtype=torch.float32, device=f"cuda:{device}")
    if cols:
        t = t.reshape(-1, cols)
    return t.contiguous()


def _stats_ptr(stats):
    if stats is None:
        return None
    if stats.dtype.itemsize != 8 or stats.numel() < 5 or not stats.is_cuda or not stats.is_contiguous():
        raise ValueError("stats must be a contiguous int64 device tensor of at least 5 counters")
    return stats.data_ptr()


def _need_rows(t, n, name):
    """Companion arrays must hold one row per query: the kernels read row i of
    every input for query i."""
    if t is not None and t.shape[0] != n:
        raise ValueError(f"{name} has {t.shape[0]} rows for {n} queries")


def _check_qoff(q_off, n_meshes, n):
    """q_off: n_meshes + 1 non-decreasing offsets from 0 to the number of
    queries (the library reads n_meshes + 1 host entries and writes outputs
    for queries [0, q_off[-1]))."""
    if q_off.ndim != 1 or q_off.shape[0] != n_meshes + 1:
        raise ValueError(f"q_off must hold n_meshes + 1 = {n_meshes + 1} offsets, got {q_off.shape}")
    if q_off[0] != 0 or np.any(np.diff(q_off.astype(np.int64)) < 0):
        raise ValueError("q_off must start at 0 and must not decrease")
    if int(q_off[-1]) != n:
        raise ValueError(f"q_off[-1] = {int(q_off[-1])} but there are {n} queries")


def _order_ptr(order, n):
    if order is None:
        return None
    if order.dtype.itemsize != 4 or order.numel() != n or not order.is_cuda or not order.is_contiguous():
        raise ValueError("order must be a contiguous 32-bit device tensor with one entry per query")
    return order.data_ptr()


class Scene:
    """``lbvh::scene<3>`` on one GPU (scene.cuh:705-1268)."""

    dim = 3

    def __init__(self, vertices, indices, device=0):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("snch_lbvh_amd needs a ROCm GPU; none is visible")
        self.device = device
        torch.cuda.set_device(device)
        d = self.dim
        V = np.ascontiguousarray(np.asarray(vertices, np.float32).reshape(-1, d))
        F = np.ascontiguousarray(np.asarray(indices, np.int32).reshape(-1, d))
        self.V, self.F = V, F
        h = ctypes.c_void_p()
        create = lib().snch_scene_create_tri if d == 3 else lib().snch_scene_create_seg
        check(create(context(device), V.ctypes.data, V.shape[0], F.ctypes.data, F.shape[0], ctypes.byref(h)))
        self._h = h
        dim, n_prims, n_sil = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().snch_scene_info(h, ctypes.byref(dim), ctypes.byref(n_prims), ctypes.byref(n_sil)))
        self.n_prims, self.n_silhouettes = n_prims.value, n_sil.value
        self.built = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.snch_scene_destroy(h)
            self._h = None

    # scene<3>::compute_silhouettes (scene.cuh:1167-1204) runs at construction
    def compute_silhouettes(self):
        return self

    def build_bvh(self, reference_radius=False, leaf_size=1, time_phases=False, top_layout="auto"):
        """scene<3>::build_bvh (scene.cuh:1205-1229) -> bvh::construct on the GPU.
        time_phases: also record the GPU time of each build phase (build_phases).
        top_layout: "auto", or "memory" / "climb" to force the layout of the
        refit above 16384-leaf subtrees (the tree is the same)."""
        flags = (BUILD_REFERENCE_RADIUS if reference_radius else 0) | (BUILD_TIME_PHASES if time_phases else 0)
        flags |= _TOP_LAYOUT[top_layout]
        opts = BuildOpts(flags, leaf_size)
        check(lib().snch_scene_build(self._h, ctypes.byref(opts), _stream()))
        self.built = True
        return self

    def build_ms(self):
        ms = ctypes.c_float()
        check(lib().snch_scene_build_time(self._h, ctypes.byref(ms)))
        return ms.value

    def build_phases(self):
        """GPU milliseconds of the last build's phases (zeros unless it was
        built with time_phases=True): leaf generation + Morton codes, radix
        sort, Karras emission, refit."""
        ms = (ctypes.c_float * 4)()
        check(lib().snch_scene_build_phases(self._h, ms))
        return {"leaves_ms": ms[0], "sort_ms": ms[1], "karras_ms": ms[2], "refit_ms": ms[3]}

    def query_order(self, q):
        """Morton visiting order of the query points (int32 device tensor)."""
        import torch

        q = _dev_f32(q, self.dim, self.device)
        perm = torch.empty(q.shape[0], dtype=torch.int32, device=q.device)
        check(lib().snch_query_order(self._h, _ptr(q), q.shape[0], _ptr(perm), _stream()))
        return perm

    def closest_point(self, q, stats=None, coherent=0, order=None, block_size=0, stack_depth=0):
        """stats: optional int64 CUDA tensor of 5 counters (child box tests,
        leaf evaluations, primitive tests, internal nodes visited, cone
        records read), accumulated."""
        import torch

        q = _dev_f32(q, self.dim, self.device)
        n = q.shape[0]
        idx = torch.empty(n, dtype=torch.int32, device=q.device)
        dist = torch.empty(n, dtype=torch.float32, device=q.device)
        opts = QueryOpts(block_size, coherent, _stats_ptr(stats), _order_ptr(order, n), stack_depth, 0)
        check(lib().snch_closest_point_ex(self._h, _ptr(q), n, _ptr(idx), _ptr(dist), ctypes.byref(opts), _stream()))
        return idx, dist

    def closest_silhouette(self, q, flip=False, seed_radius=None, block_size=0, stats=None, coherent=0, order=None,
                           stack_depth=0, seed_is_cap=False, seed_in_place=False):
        """query_device(nearest_silhouette) (query.cuh:325-423) over q (n x dim).
        seed_radius: optional per-query pruning radius (a learned bound); the
        outputs equal the unseeded call's.  Queries not settled within it are
        re-queued for whole wavefronts of the same launch, or with
        seed_in_place searched again at once by their own lane.  seed_is_cap
        makes the radius a hard search radius."""
        import torch

        q = _dev_f32(q, self.dim, self.device)
        n = q.shape[0]
        seed = _dev_f32(seed_radius, 0, self.device).reshape(-1) if seed_radius is not None else None
        _need_rows(seed, n, "seed_radius")
        sil = torch.empty(n, dtype=torch.int32, device=q.device)
        dist = torch.empty(n, dtype=torch.float32, device=q.device)
        opts = QueryOpts(block_size, coherent, _stats_ptr(stats), _order_ptr(order, n), stack_depth, 0,
                         int(bool(seed_is_cap)), int(bool(seed_in_place)))
        check(lib().snch_closest_silhouette_ex(self._h, _ptr(q), n, int(bool(flip)), _ptr(seed), _ptr(sil),
                                               _ptr(dist), ctypes.byref(opts), _stream()))
        return sil, dist

    def ray_intersect(self, o, d, tmax=None, any_hit=False, stats=None):
        import torch

        o = _dev_f32(o, self.dim, self.device)
        d = _dev_f32(d, self.dim, self.device)
        n = o.shape[0]
        tm = _dev_f32(tmax, 0, self.device).reshape(-1) if tmax is not None else None
        _need_rows(d, n, "d")
        _need_rows(tm, n, "tmax")
        hit = torch.empty(n, dtype=torch.uint8, device=o.device)
        t = torch.empty(n, dtype=torch.float32, device=o.device)
        uv = torch.empty((n, 2) if self.dim == 3 else (n,), dtype=torch.float32, device=o.device)
        idx = torch.empty(n, dtype=torch.int32, device=o.device)
        opts = QueryOpts(0, 0, _stats_ptr(stats), None, 0, 0)
        check(lib().snch_ray_ex(self._h, _ptr(o), _ptr(d), _ptr(tm), n, int(bool(any_hit)), _ptr(hit), _ptr(t),
                                _ptr(uv), _ptr(idx), ctypes.byref(opts), _stream()))
        return hit.bool(), t, uv, idx

    def sample_in_sphere(self, c, r, u, with_point=True):
        import torch

        c = _dev_f32(c, self.dim, self.device)
        n = c.shape[0]
        r = _dev_f32(r, 0, self.device).reshape(-1)
        u = _dev_f32(u, 3, self.device)
        _need_rows(r, n, "r")
        _need_rows(u, n, "u")
        idx = torch.empty(n, dtype=torch.int32, device=c.device)
        pdf = torch.empty(n, dtype=torch.float32, device=c.device)
        pt = torch.empty((n, self.dim), dtype=torch.float32, device=c.device) if with_point else None
        check(lib().snch_sample_in_sphere(self._h, _ptr(c), _ptr(r), _ptr(u), n, _ptr(idx), _ptr(pdf), _ptr(pt),
                                          _stream()))
        return idx, pdf, pt

    def sample_on_object(self, idx, uv):
        import torch

        idx = torch.as_tensor(idx, dtype=torch.int32, device=f"cuda:{self.device}").contiguous()
        uv = _dev_f32(uv, 2, self.device)
        _need_rows(uv, idx.shape[0], "uv")
        pt = torch.empty((idx.shape[0], self.dim), dtype=torch.float32, device=idx.device)
        check(lib().snch_sample_on_object(self._h, _ptr(idx), _ptr(uv), idx.shape[0], _ptr(pt), _stream()))
        return pt

    def overlap(self, lo, hi, max_per_query=64):
        """query_device(bvh, overlaps(box)) (query.cuh:176-231) for boxes
        [lo[i], hi[i]]: (count per query, ids: the first max_per_query in the
        reference's visit order, -1 beyond the count)."""
        import torch

        lo = _dev_f32(lo, self.dim, self.device)
        hi = _dev_f32(hi, self.dim, self.device)
        n = lo.shape[0]
        _need_rows(hi, n, "hi")
        cnt = torch.empty(n, dtype=torch.int32, device=lo.device)
        idx = torch.full((n, max_per_query), -1, dtype=torch.int32, device=lo.device)
        check(lib().snch_overlap(self._h, _ptr(lo), _ptr(hi), n, max_per_query, _ptr(cnt),
                                 _ptr(idx) if max_per_query else None, _stream()))
        return cnt, idx

    def line_intersect(self, o, d, max_per_query=16):
        """query_device(bvh, line_intersect(line(o, d))) (query.cuh:12-77):
        (count per query, ids, line parameters t), ids in the reference's
        visit order, -1 / nan beyond the count."""
        import torch

        o = _dev_f32(o, self.dim, self.device)
        d = _dev_f32(d, self.dim, self.device)
        n = o.shape[0]
        _need_rows(d, n, "d")
        cnt = torch.empty(n, dtype=torch.int32, device=o.device)
        idx = torch.full((n, max_per_query), -1, dtype=torch.int32, device=o.device)
        t = torch.full((n, max_per_query), float("nan"), dtype=torch.float32, device=o.device)
        check(lib().snch_line_intersect(self._h, _ptr(o), _ptr(d), n, max_per_query, _ptr(cnt),
                                        _ptr(idx) if max_per_query else None, _ptr(t) if max_per_query else None,
                                        _stream()))
        return cnt, idx, t

    def device_view(self):
        """snch_scene_device_view: the built tree for users' own HIP kernels
        (include/snch_lbvh/device.hpp), a DeviceView of device pointers."""
        v = DeviceView()
        check(lib().snch_scene_device_view(self._h, 0, ctypes.byref(v)))
        return v

    def export(self):
        """The tree in the reference layout (host numpy arrays)."""
        N = self.n_prims
        NN = 2 * N - 1
        nodes = np.empty((NN, 4), np.uint32)
        boxes = np.empty((NN, 6), np.float32)
        cones = np.empty((NN, 5), np.float32)
        codes = np.empty(N, np.uint32)
        sidx = np.empty(N, np.uint32)
        check(lib().snch_scene_export(self._h, nodes.ctypes.data, boxes.ctypes.data, cones.ctypes.data,
                                      codes.ctypes.data, sidx.ctypes.data))
        return dict(nodes=nodes, boxes=boxes, cones=cones, sorted_codes=codes, sorted_idx=sidx)

    def silhouettes(self):
        slots = np.empty((self.n_silhouettes, 4), np.int32)
        owned = np.empty((self.n_prims, self.dim), np.int32)
        check(lib().snch_scene_silhouettes(self._h, slots.ctypes.data, owned.ctypes.data))
        return slots, owned


class Scene2D(Scene):
    """``lbvh::scene<2>``: line segments with silhouette vertices (scene.cuh:287-703)."""

    dim = 2


class BatchScene:
    """Several independent triangle meshes built and queried together: one
    build and one persistent query launch cover every mesh.  No counterpart
    in the reference (one lbvh::scene<3> per mesh there)."""

    dim = 3

    def __init__(self, meshes, device=0):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("snch_lbvh_amd needs a ROCm GPU; none is visible")
        self.device = device
        torch.cuda.set_device(device)
        Vs = [np.ascontiguousarray(np.asarray(v, np.float32).reshape(-1, 3)) for v, _ in meshes]
        Fs = [np.ascontiguousarray(np.asarray(f, np.int32).reshape(-1, 3)) for _, f in meshes]
        self.meshes = list(zip(Vs, Fs))
        v_off = np.zeros(len(Vs) + 1, np.uint32)
        f_off = np.zeros(len(Fs) + 1, np.uint32)
        v_off[1:] = np.cumsum([v.shape[0] for v in Vs])
        f_off[1:] = np.cumsum([f.shape[0] for f in Fs])
        V = np.ascontiguousarray(np.concatenate(Vs)) if Vs else np.zeros((0, 3), np.float32)
        F = np.ascontiguousarray(np.concatenate(Fs)) if Fs else np.zeros((0, 3), np.int32)
        self.f_off = f_off
        h = ctypes.c_void_p()
        check(lib().snch_scene_create_tri_batch(context(device), len(Vs), V.ctypes.data, v_off.ctypes.data,
                                                F.ctypes.data, f_off.ctypes.data, ctypes.byref(h)))
        self._h = h
        self.n_meshes = len(Vs)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.snch_scene_destroy(h)
            self._h = None

    def build_bvh(self, reference_radius=False, leaf_size=1, time_phases=False, top_layout="auto"):
        flags = (BUILD_REFERENCE_RADIUS if reference_radius else 0) | (BUILD_TIME_PHASES if time_phases else 0)
        flags |= _TOP_LAYOUT[top_layout]
        opts = BuildOpts(flags, leaf_size)
        check(lib().snch_scene_build(self._h, ctypes.byref(opts), _stream()))
        return self

    def build_ms(self):
        ms = ctypes.c_float()
        check(lib().snch_scene_build_time(self._h, ctypes.byref(ms)))
        return ms.value

    def build_phases(self):
        """GPU milliseconds of the last build's phases (see Scene.build_phases)."""
        ms = (ctypes.c_float * 4)()
        check(lib().snch_scene_build_phases(self._h, ms))
        return {"leaves_ms": ms[0], "sort_ms": ms[1], "karras_ms": ms[2], "refit_ms": ms[3]}

    @staticmethod
    def _qcat(queries):
        q_off = np.zeros(len(queries) + 1, np.uint32)
        q_off[1:] = np.cumsum([len(x) for x in queries])
        return q_off

    def _opts(self, tune, coherent):
        t = tune or {}
        return QueryOpts(t.get("block_size", 0), coherent, None, None, t.get("stack_depth", 0),
                         t.get("queries_per_wave", 0))

    def closest_point(self, q, q_off, tune=None, coherent=0):
        """q: (n, 3) points of all meshes, grouped by mesh; q_off: n_meshes + 1
        offsets.  Returns (index local to the mesh, distance) per query."""
        import torch

        q = _dev_f32(q, 3, self.device)
        q_off = np.ascontiguousarray(q_off, np.uint32)
        n = q.shape[0]
        _check_qoff(q_off, self.n_meshes, n)
        idx = torch.empty(n, dtype=torch.int32, device=q.device)
        dist = torch.empty(n, dtype=torch.float32, device=q.device)
        opts = self._opts(tune, coherent)
        check(lib().snch_batch_closest_point(self._h, _ptr(q), q_off.ctypes.data, _ptr(idx), _ptr(dist),
                                             ctypes.byref(opts), _stream()))
        return idx, dist

    def closest_silhouette(self, q, q_off, flip=False, tune=None, coherent=0, seed_radius=None):
        """seed_radius: optional per-query pruning radius (a learned bound);
        results equal the unseeded call's."""
        import torch

        q = _dev_f32(q, 3, self.device)
        q_off = np.ascontiguousarray(q_off, np.uint32)
        n = q.shape[0]
        _check_qoff(q_off, self.n_meshes, n)
        seed = _dev_f32(seed_radius, 0, self.device).reshape(-1) if seed_radius is not None else None
        _need_rows(seed, n, "seed_radius")
        sil = torch.empty(n, dtype=torch.int32, device=q.device)
        dist = torch.empty(n, dtype=torch.float32, device=q.device)
        opts = self._opts(tune, coherent)
        check(lib().snch_batch_closest_silhouette(self._h, _ptr(q), q_off.ctypes.data, int(bool(flip)), _ptr(seed),
                                                  _ptr(sil), _ptr(dist), ctypes.byref(opts), _stream()))
        return sil, dist

    def autotune(self, q, q_off, kind="silhouette", sample_per_mesh=0, flip=False):
        """Sweeps leaf size, block size, stack depth and queries per wave;
        rebuilds with the best leaf size and returns the choice as a dict."""
        q = _dev_f32(q, 3, self.device)
        q_off = np.ascontiguousarray(q_off, np.uint32)
        _check_qoff(q_off, self.n_meshes, q.shape[0])
        r = TuneResult()
        check(lib().snch_batch_autotune(self._h, 1 if kind == "silhouette" else 0, _ptr(q), q_off.ctypes.data,
                                        sample_per_mesh, int(bool(flip)), ctypes.byref(r), _stream()))
        return {k: getattr(r, k) for k, _ in TuneResult._fields_}

    def device_view(self, mesh):
        """snch_scene_device_view for mesh `mesh` of the batch."""
        v = DeviceView()
        check(lib().snch_scene_device_view(self._h, mesh, ctypes.byref(v)))
        return v

    def export(self, mesh):
        V, F = self.meshes[mesh]
        N = F.shape[0]
        nodes = np.empty((2 * N - 1, 4), np.uint32)
        boxes = np.empty((2 * N - 1, 6), np.float32)
        cones = np.empty((2 * N - 1, 5), np.float32)
        codes = np.empty(N, np.uint32)
        sidx = np.empty(N, np.uint32)
        check(lib().snch_scene_export_mesh(self._h, mesh, nodes.ctypes.data, boxes.ctypes.data, cones.ctypes.data,
                                           codes.ctypes.data, sidx.ctypes.data))
        return {"nodes": nodes, "boxes": boxes, "cones": cones, "sorted_codes": codes, "sorted_idx": sidx}


class LearnedBound:
    """Fourier-feature MLP (BF16 MFMA) that predicts a closest-silhouette
    distance and turns it into a seed radius for Scene.closest_silhouette.
    No counterpart in the reference; it never changes a query result."""

    def __init__(self, lo, hi, sigma=4.0, seed=0, device=0):
        self.device = device
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        h = ctypes.c_void_p()
        check(lib().snch_mlp_create(context(device), lo.ctypes.data, hi.ctypes.data, float(sigma), int(seed),
                                    ctypes.byref(h)))
        self._h = h
        self.n_params = lib().snch_mlp_num_params()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.snch_mlp_destroy(h)
            self._h = None

    def params(self):
        th = np.empty(self.n_params, np.float32)
        four = np.empty(96, np.float32)
        frame = np.empty(8, np.float32)
        check(lib().snch_mlp_params(self._h, th.ctypes.data, None, four.ctypes.data, frame.ctypes.data))
        return th, four, frame

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        assert th.size == self.n_params
        check(lib().snch_mlp_params(self._h, None, th.ctypes.data, None, None))

    def set_loss(self, tau=0.5):
        """Expectile loss weight: tau > 0.5 penalises under-prediction more,
        training an upper bound on the distance."""
        check(lib().snch_mlp_set_loss(self._h, float(tau)))
        return self

    def train(self, pts, dist, batch=16384, epochs=1, lr=1e-3):
        p = _dev_f32(pts, 3, self.device)
        d = _dev_f32(dist, 0, self.device).reshape(-1)
        losses = np.zeros(epochs, np.float32)
        check(lib().snch_mlp_train(self._h, _ptr(p), _ptr(d), p.shape[0], batch, epochs, float(lr),
                                   losses.ctypes.data, _stream()))
        return losses

    def gradient(self, pts, dist):
        p = _dev_f32(pts, 3, self.device)
        d = _dev_f32(dist, 0, self.device).reshape(-1)
        g = np.empty(self.n_params, np.float32)
        loss = ctypes.c_float()
        check(lib().snch_mlp_gradient(self._h, _ptr(p), _ptr(d), p.shape[0], g.ctypes.data, ctypes.byref(loss),
                                      _stream()))
        return g, loss.value

    def predict(self, pts, out=None):
        import torch

        p = _dev_f32(pts, 3, self.device)
        out = torch.empty(p.shape[0], dtype=torch.float32, device=p.device) if out is None else out
        check(lib().snch_mlp_predict(self._h, _ptr(p), p.shape[0], 0, 0.0, 0.0, _ptr(out), _stream()))
        return out

    def seed_radius(self, pts, rel=0.1, abs_margin=0.002, out=None):
        import torch

        p = _dev_f32(pts, 3, self.device)
        out = torch.empty(p.shape[0], dtype=torch.float32, device=p.device) if out is None else out
        check(lib().snch_mlp_predict(self._h, _ptr(p), p.shape[0], 1, float(rel), float(abs_margin), _ptr(out),
                                     _stream()))
        return out


class WoSt:
    """Walk on stars for the Laplace equation: Dirichlet data g = x y on the
    ``dirichlet`` scene, zero Neumann flux on the ``neumann`` scene (or none).
    Each step runs closest point, closest silhouette, ray and
    SampleTriangleInSphere over the live walks; the loop can run as a
    captured hipGraph."""

    def __init__(self, dirichlet, neumann=None, max_walks=1 << 20):
        self.device = dirichlet.device
        self.max_walks = max_walks
        self._scenes = (dirichlet, neumann)  # keep the scenes alive
        h = ctypes.c_void_p()
        check(lib().snch_wost_create(context(self.device), dirichlet._h, neumann._h if neumann else None, max_walks,
                                     ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.snch_wost_destroy(h)
            self._h = None

    def solve(self, x0, max_steps=64, eps=1e-3, seed=0, use_graph=True, out=None):
        import torch

        x = _dev_f32(x0, 3, self.device)
        n = x.shape[0]
        out = torch.empty(n, dtype=torch.float32, device=x.device) if out is None else out
        live = ctypes.c_uint32()
        check(lib().snch_wost_solve(self._h, _ptr(x), n, max_steps, float(eps), int(seed), int(bool(use_graph)),
                                    _ptr(out), ctypes.byref(live), _stream()))
        return out, live.value
