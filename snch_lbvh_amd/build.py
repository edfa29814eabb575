"""Compile libsnch.so (HIP kernels + C ABI) for gfx950 with hipcc.

Run ``python -m snch_lbvh_amd.build`` (or ``__graft_entry__.build()``).  The
library lands next to this file as ``libsnch.so``.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsnch.so")
BUILD_DIR = os.path.join(HERE, "_build")
SOURCES = ["capi.hip", "build.hip", "sort.hip", "edges.hip", "query.hip", "mlp.hip", "wost.hip", "noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SNCH_OFFLOAD_ARCH", "gfx950")
# -ffp-contract=off: results are compared bit for bit with IEEE single-precision
# CPU evaluation of the same expressions.  Packed FP32 instructions are kept out.
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
]


def _compile(src):
    obj = os.path.join(BUILD_DIR, src + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


def build_device_api_test(verbose=False):
    """Compile tests/cpp/device_api.hip (a user's kernel on the device header
    include/snch_lbvh/device.hpp) into tests/cpp/libdevice_api.so, linked
    against libsnch.so."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "device_api.hip")
    out = os.path.join(root, "tests", "cpp", "libdevice_api.so")
    if not os.path.exists(OUT):
        build(verbose)
    cmd = [HIPCC, *FLAGS, "-shared", "-I", os.path.join(root, "include"), src, "-L", HERE, "-lsnch",
           "-Wl,-rpath,$ORIGIN/../../snch_lbvh_amd", "-o", out + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for device_api.hip:\n{r.stdout}\n{r.stderr}")
    os.replace(out + ".tmp", out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
