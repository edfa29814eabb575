"""CPU-only checks of the C ABI: the library loads, exports every entry point
that include/snch.h declares, and reports errors without a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for h in ("include/snch.h",):
        txt = open(os.path.join(ROOT, h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"\b(snch_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


@pytest.fixture(scope="module")
def libsnch():
    import snch_lbvh_amd as S

    if not os.path.exists(S.LIB_PATH):
        from snch_lbvh_amd import build

        build.build()
    return S.lib()


def test_exports_every_declared_symbol(libsnch):
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(libsnch, n), n


def test_python_binding_declares_all(libsnch):
    import snch_lbvh_amd as S

    assert sorted(S.EXPORTED) == _declared()


def test_errors_without_gpu(libsnch):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    c = ctypes.c_void_p()
    rc = libsnch.snch_context_create(0, ctypes.byref(c))
    assert rc != 0
    assert libsnch.snch_last_error()
    # null-argument handling needs no device
    assert libsnch.snch_scene_info(None, None, None, None) == 1
    assert libsnch.snch_closest_point(None, None, 0, None, None, None) == 1
