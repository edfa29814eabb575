"""The HIP device API for users' own kernels (include/snch_lbvh/device.hpp):
a user kernel (tests/cpp/device_api.hip, built by __graft_entry__.build())
calls every per-thread query, one thread per query, on a scene's device view;
its answers must equal the library's batched queries bit for bit (the 2D
ray's hits past a segment end aside, which depend on the visiting order in
the reference as well)."""
import ctypes
import os

import numpy as np
import pytest

from snch_lbvh_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "tests", "cpp", "libdevice_api.so")


def _lib():
    import snch_lbvh_amd as S

    S.lib()  # libsnch.so first: the test library links to the same copy
    if not os.path.exists(SO):
        from snch_lbvh_amd import build

        build.build_device_api_test()
    L = ctypes.CDLL(SO)
    vp = ctypes.c_void_p
    L.run_device_api.argtypes = [vp, ctypes.c_uint32] + [vp] * 5 + [ctypes.c_uint32, ctypes.c_int] + [vp] * 13
    L.run_device_api.restype = ctypes.c_int
    return L


def test_device_api_library_loads():
    """The user-kernel library builds against the public header and exports its
    entry point; libsnch.so exports snch_scene_device_view (no GPU needed)."""
    import snch_lbvh_amd as S

    L = _lib()
    assert hasattr(L, "run_device_api")
    assert hasattr(S.lib(), "snch_scene_device_view")


def _inputs(V, dim, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = W.padded_box(V, 0.25)
    q = rng.uniform(lo, hi, (n, dim)).astype(np.float32)
    d = rng.normal(size=(n, dim))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    tmax = np.where(np.arange(n) % 4 == 0, np.float32(0.3), np.float32(np.inf)).astype(np.float32)
    r = W.log_uniform(n, 1e-3, 0.5, seed + 1)
    u = rng.random((n, 3)).astype(np.float32)
    return q, d, tmax, r, u


def _run(L, S, sc, mesh, q, d, tmax, r, u, flip):
    import torch

    n = q.shape[0]
    dev = torch.device("cuda")
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in
         (("q", q), ("d", d), ("tmax", tmax), ("r", r), ("u", u))}
    out = {
        "cp_idx": torch.empty(n, dtype=torch.int32, device=dev), "cp_dist": torch.empty(n, device=dev),
        "sil_idx": torch.empty(n, dtype=torch.int32, device=dev), "sil_dist": torch.empty(n, device=dev),
        "hit": torch.empty(n, dtype=torch.uint8, device=dev), "t": torch.empty(n, device=dev),
        "uv": torch.empty(2 * n, device=dev), "ray_idx": torch.empty(n, dtype=torch.int32, device=dev),
        "any_hit": torch.empty(n, dtype=torch.uint8, device=dev),
        "smp_idx": torch.empty(n, dtype=torch.int32, device=dev), "smp_pdf": torch.empty(n, device=dev),
        "smp_pt": torch.empty(3 * n, device=dev),
    }
    p = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
    rc = L.run_device_api(sc._h, mesh, p(t["q"]), p(t["d"]), p(t["tmax"]), p(t["r"]), p(t["u"]), n, int(flip),
                          *(p(out[k]) for k in ("cp_idx", "cp_dist", "sil_idx", "sil_dist", "hit", "t", "uv",
                                                "ray_idx", "any_hit", "smp_idx", "smp_pdf", "smp_pt")),
                          S._stream())
    assert rc == 0, rc
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["torus", "ico_leaf4", "circle_square"])
def test_device_api_equals_batched(case):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import snch_lbvh_amd as S

    L = _lib()
    if case == "circle_square":
        from snch_lbvh_amd.scene_loader import load_obj

        V, P = load_obj(os.path.join(ROOT, "tests", "golden", "circle_square.obj"), 2)
        sc = S.Scene2D(V, P).build_bvh()
        dim = 2
    else:
        V, P = W.cut_torus(96, 48) if case == "torus" else W.icosphere(4, 0.02, seed=2)
        sc = S.Scene(V, P).build_bvh(leaf_size=4 if case == "ico_leaf4" else 1)
        dim = 3
    q, d, tmax, r, u = _inputs(V, dim, 20000, 3)
    for flip in (False, True):
        g = _run(L, S, sc, 0, q, d, tmax, r, u, flip)
        ci, cd = (x.cpu().numpy() for x in sc.closest_point(q))
        np.testing.assert_array_equal(g["cp_dist"], cd)
        np.testing.assert_array_equal(g["cp_idx"], ci)
        si, sd = (x.cpu().numpy() for x in sc.closest_silhouette(q, flip=flip))
        np.testing.assert_array_equal(g["sil_dist"], sd)
        np.testing.assert_array_equal(g["sil_idx"], si)
    hit, t, uv, idx = (x.cpu().numpy() for x in sc.ray_intersect(q, d, tmax))
    gh = g["hit"].astype(bool)
    if dim == 3:
        np.testing.assert_array_equal(gh, hit)
        np.testing.assert_array_equal(g["t"][hit], t[hit])
        np.testing.assert_array_equal(g["ray_idx"][hit], idx[hit])
        np.testing.assert_array_equal(g["uv"].reshape(-1, 2)[hit], uv.reshape(-1, 2)[hit])
    else:
        assert (gh == hit).mean() > 0.99
    anyh = sc.ray_intersect(q, d, tmax, any_hit=True)[0].cpu().numpy()
    np.testing.assert_array_equal(g["any_hit"].astype(bool), anyh)
    si, sp, spt = (x.cpu().numpy() for x in sc.sample_in_sphere(q, r, u))
    np.testing.assert_array_equal(g["smp_idx"], si)
    np.testing.assert_array_equal(g["smp_pdf"], sp)
    np.testing.assert_array_equal(g["smp_pt"].reshape(-1, 3)[:, :dim], spt.reshape(-1, dim))


@pytest.mark.gpu
def test_device_api_batch_mesh():
    """A batch's per-mesh views: mesh-local ids, equal to the batched queries."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import snch_lbvh_amd as S

    L = _lib()
    meshes = [W.icosphere(3, 0.02, seed=5), W.cut_torus(64, 40, seed=6), W.icosphere(2, 0.0)]
    b = S.BatchScene(meshes).build_bvh()
    for m, (V, _) in enumerate(meshes):
        q, d, tmax, r, u = _inputs(V, 3, 3000, 10 + m)
        g = _run(L, S, b, m, q, d, tmax, r, u, False)
        # the batched call with queries for mesh m only
        q_off = np.array([0] * (m + 1) + [len(q)] * (len(meshes) - m), np.uint32)
        qq = q
        ci, cd = (x.cpu().numpy() for x in b.closest_point(qq, q_off))
        np.testing.assert_array_equal(g["cp_dist"], cd)
        np.testing.assert_array_equal(g["cp_idx"], ci)
        si, sd = (x.cpu().numpy() for x in b.closest_silhouette(qq, q_off))
        np.testing.assert_array_equal(g["sil_dist"], sd)
        np.testing.assert_array_equal(g["sil_idx"], si)
