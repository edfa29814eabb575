"""The header-only C++ front end (include/snch_lbvh/scene.hpp) compiles with a
plain host compiler against libsnch.so, and (on a GPU) runs end to end."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _compile(tmp_path):
    from snch_lbvh_amd import LIB_PATH, build

    if not os.path.exists(LIB_PATH):
        build.build()
    exe = str(tmp_path / "scene_api")
    libdir = os.path.dirname(LIB_PATH)
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "cpp", "scene_api.cpp"), "-L", libdir, "-lsnch",
                           f"-Wl,-rpath,{libdir}", "-o", exe])
    return exe


def test_cpp_front_end_compiles(tmp_path):
    assert os.path.exists(_compile(tmp_path))


@pytest.mark.gpu
def test_cpp_front_end_runs(tmp_path):
    exe = _compile(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok")
