"""The OBJ loaders (Python snch_lbvh_amd.scene_loader and the C++
include/snch_lbvh/scene_loader.hpp) against the fixtures' expected arrays,
written from the generating arrays (tests/golden/make_obj_fixtures.py), and
their error behaviour (scene_loader.cuh:13-152)."""
import json
import os
import subprocess

import numpy as np
import pytest

from snch_lbvh_amd.scene_loader import SceneLoader, load_obj

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXPECTED = json.load(open(os.path.join(GOLD, "obj_expected.json")))


@pytest.fixture(scope="module")
def obj_dump(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("obj") / "obj_dump")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "cpp", "obj_dump.cpp"), "-o", exe])
    return exe


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_python_loader(name):
    e = EXPECTED[name]
    ld = SceneLoader(os.path.join(GOLD, name), e["dim"])
    np.testing.assert_array_equal(ld.get_vertices(), np.float32(e["vertices"]))
    np.testing.assert_array_equal(ld.get_indices(), np.int32(e["indices"]))
    assert ld.vertices_size() == len(e["vertices"]) and ld.primitives_size() == len(e["indices"])


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_cpp_loader(obj_dump, name):
    e = EXPECTED[name]
    r = subprocess.run([obj_dump, str(e["dim"]), os.path.join(GOLD, name)], capture_output=True, text=True,
                       check=True)
    lines = r.stdout.split("\n")
    nv, npr = (int(x) for x in lines[0].split())
    d = e["dim"]
    vals = lines[1:1 + nv * d]
    idx = lines[1 + nv * d:1 + nv * d + npr * d]
    np.testing.assert_array_equal(np.float32([float(x) for x in vals]).reshape(-1, d), np.float32(e["vertices"]))
    np.testing.assert_array_equal(np.int32([int(x) for x in idx]).reshape(-1, d), np.int32(e["indices"]))


def test_errors(obj_dump, tmp_path):
    with pytest.raises(OSError, match="Could not open .obj file."):
        load_obj(str(tmp_path / "missing.obj"))
    r = subprocess.run([obj_dump, "3", str(tmp_path / "missing.obj")], capture_output=True, text=True)
    assert r.returncode == 3 and "Could not open .obj file." in r.stdout
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2\n")
    with pytest.raises(ValueError):
        load_obj(str(bad))
    r = subprocess.run([obj_dump, "3", str(bad)], capture_output=True, text=True)
    assert r.returncode == 3 and "three vertex indices" in r.stdout
    bad.write_text("v 0 0\n")
    with pytest.raises(ValueError):
        load_obj(str(bad))
    r = subprocess.run([obj_dump, "3", str(bad)], capture_output=True, text=True)
    assert r.returncode == 3
