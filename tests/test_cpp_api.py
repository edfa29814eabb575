"""The header-only C++ front end (include/snch_lbvh/lbvh.hpp: scene<2|3>,
the query descriptors, scene_loader) compiles with a plain host compiler
against libsnch.so, and so does the C++ example in INTEGRATION.md (extracted
from the document, so the two cannot drift apart).  On a GPU the program
drives scene<3> and scene<2> from OBJ fixtures through every descriptor and
its outputs equal the oracle's."""
import os
import re
import subprocess

import numpy as np
import pytest

from snch_lbvh_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _compile(tmp_path, src, name):
    from snch_lbvh_amd import LIB_PATH, build

    if not os.path.exists(LIB_PATH):
        build.build()
    exe = str(tmp_path / name)
    libdir = os.path.dirname(LIB_PATH)
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), src, "-L",
                           libdir, "-lsnch", f"-Wl,-rpath,{libdir}", "-o", exe])
    return exe


def _integration_snippet(tmp_path):
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", txt, flags=re.S)
    assert blocks, "INTEGRATION.md has no C++ example"
    src = tmp_path / "integration_example.cpp"
    src.write_text(blocks[0])
    return str(src)


def test_cpp_front_end_compiles(tmp_path):
    assert os.path.exists(_compile(tmp_path, os.path.join(ROOT, "tests", "cpp", "scene_api.cpp"), "scene_api"))


def test_integration_example_compiles(tmp_path):
    assert os.path.exists(_compile(tmp_path, _integration_snippet(tmp_path), "integration_example"))


def _inputs(V, dim, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = V.min(0), V.max(0)
    ext = hi - lo
    pts = rng.uniform(lo - 0.2 * ext, hi + 0.2 * ext, (n, dim)).astype(np.float32)
    d = rng.normal(size=(n, dim))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    tmax = np.where(np.arange(n) % 3 == 0, np.float32(0.4), np.float32(np.inf)).astype(np.float32)
    r = W.log_uniform(n, 1e-3, 0.5, seed + 1)
    u = rng.random((n, 3)).astype(np.float32)
    e = (rng.random((n, dim)) * 0.2).astype(np.float32)
    return {"points": pts, "dirs": d, "tmax": tmax, "radii": r, "u3": u, "lo": pts - e, "hi": pts + e}


@pytest.mark.gpu
def test_cpp_front_end_runs(tmp_path, oracle):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from snch_lbvh_amd.scene_loader import load_obj

    exe = _compile(tmp_path, os.path.join(ROOT, "tests", "cpp", "scene_api.cpp"), "scene_api")
    V3, F3 = load_obj(os.path.join(GOLD, "ico2.obj"), 3)
    V2, S2 = load_obj(os.path.join(GOLD, "circle_square.obj"), 2)
    n = 3000
    ins = {"3": _inputs(V3, 3, n, 1), "2": _inputs(V2, 2, n, 2)}
    for tag, d in ins.items():
        for k, a in d.items():
            a.tofile(str(tmp_path / f"in{tag}_{k}.bin"))
    r = subprocess.run([exe, os.path.join(GOLD, "ico2.obj"), os.path.join(GOLD, "circle_square.obj"), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr

    def out(tag, name, dt):
        return np.fromfile(str(tmp_path / f"out{tag}_{name}.bin"), dtype=dt)

    t3 = oracle.build_tree(V3, F3)
    t2 = oracle.build_tree2d(V2, S2)
    for tag, t, dim in (("3", t3, 3), ("2", t2, 2)):
        x = ins[tag]
        if dim == 3:
            _, od = oracle.closest_point(t, x["points"])
            _, osd = oracle.closest_silhouette(t, x["points"])
            oh, ot, _, _ = oracle.ray(t, x["points"], x["dirs"], x["tmax"])
            oi, op, opt = oracle.sample_in_sphere(t, x["points"], x["radii"], x["u3"])
        else:
            _, od = oracle.closest_point2d(t, x["points"])
            _, osd = oracle.closest_silhouette2d(t, x["points"])
            oh, ot, _, _ = oracle.ray2d(t, x["points"], x["dirs"], x["tmax"])
            oi, op, opt = oracle.sample2d(t, x["points"], x["radii"], x["u3"])
        np.testing.assert_array_equal(out(tag, "cp_dist", np.float32), od)
        np.testing.assert_array_equal(out(tag, "sil_dist", np.float32), osd)
        gh = out(tag, "ray_hit", np.uint8).astype(bool)
        if dim == 3:
            np.testing.assert_array_equal(gh, oh)
            np.testing.assert_array_equal(out(tag, "ray_t", np.float32)[oh], ot[oh])
        else:  # 2D hits past a segment end depend on the visiting order (a documented deviation)
            assert (gh == oh).mean() > 0.99
        np.testing.assert_array_equal(out(tag, "anyhit", np.uint8).astype(bool), gh)
        np.testing.assert_array_equal(out(tag, "smp_idx", np.int32), oi)
        np.testing.assert_array_equal(out(tag, "smp_pdf", np.float32), op)
        np.testing.assert_array_equal(out(tag, "smp_pt", np.float32).reshape(-1, dim), opt)
        M = 16
        oc, oix = oracle.overlap(t, x["lo"], x["hi"], M)
        np.testing.assert_array_equal(out(tag, "ov_cnt", np.uint32), oc)
        m = np.arange(M)[None, :] < np.minimum(oc, M)[:, None]
        np.testing.assert_array_equal(out(tag, "ov_idx", np.uint32).reshape(-1, M)[m], oix[m])
        lc, lix, lt = oracle.line_intersect(t, x["points"], x["dirs"], M)
        np.testing.assert_array_equal(out(tag, "ln_cnt", np.uint32), lc)
        m = np.arange(M)[None, :] < np.minimum(lc, M)[:, None]
        np.testing.assert_array_equal(out(tag, "ln_idx", np.uint32).reshape(-1, M)[m], lix[m])
        np.testing.assert_array_equal(out(tag, "ln_t", np.float32).reshape(-1, M)[m], lt[m])


@pytest.mark.gpu
def test_integration_example_runs(tmp_path):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    exe = _compile(tmp_path, _integration_snippet(tmp_path), "integration_example")
    r = subprocess.run([exe, os.path.join(GOLD, "ico2.obj")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "closest silhouette" in r.stdout
