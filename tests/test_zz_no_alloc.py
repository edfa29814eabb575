"""snch_closest_silhouette_noalloc (static device scratch, no allocation)
against snch_closest_silhouette, through the C ABI: two calls in a row, the
second smaller, both above the Morton-ordering threshold, then a small call
below it and a seeded call; sizes beyond the static scratch are rejected."""
import ctypes

import numpy as np
import pytest

from snch_lbvh_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import snch_lbvh_amd as S

    S.lib()
    return S


def _call_noalloc(S, sc, q, seed=None, flip=0):
    import torch

    n = q.shape[0]
    sil = torch.full((n,), -7, dtype=torch.int32, device=q.device)
    dist = torch.full((n,), -7.0, dtype=torch.float32, device=q.device)
    S.check(S.lib().snch_closest_silhouette_noalloc(sc._h, S._ptr(q), n, flip, S._ptr(seed), S._ptr(sil),
                                                    S._ptr(dist), S._stream()))
    return sil, dist


def _call_plain(S, sc, q, seed=None, flip=0):
    import torch

    n = q.shape[0]
    sil = torch.empty(n, dtype=torch.int32, device=q.device)
    dist = torch.empty(n, dtype=torch.float32, device=q.device)
    S.check(S.lib().snch_closest_silhouette(sc._h, S._ptr(q), n, flip, S._ptr(seed), S._ptr(sil), S._ptr(dist),
                                            S._stream()))
    return sil, dist


def test_noalloc_equals_plain(S):
    import torch

    V, F = W.cut_torus(256, 128)
    sc = S.Scene(V, F).build_bvh()
    lo, hi = W.padded_box(V)
    q1 = torch.from_numpy(W.uniform_points(1 << 20, lo, hi, seed=31)).cuda()
    q2 = torch.from_numpy(W.uniform_points(200000, lo, hi, seed=32)).cuda()
    q3 = torch.from_numpy(W.uniform_points(3000, lo, hi, seed=33)).cuda()
    for q, flip in ((q1, 0), (q2, 1), (q3, 0)):
        a = _call_noalloc(S, sc, q, flip=flip)
        b = _call_plain(S, sc, q, flip=flip)
        torch.cuda.synchronize()
        assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]), q.shape
    # seeded: same answers as unseeded
    seed = (_call_plain(S, sc, q2)[1] * 1.1).contiguous()
    seed[::7] = 0.0  # some seeds too small: those queries fall back
    a = _call_noalloc(S, sc, q2, seed=seed)
    b = _call_plain(S, sc, q2)
    torch.cuda.synchronize()
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    S.check(S.lib().snch_stream_synchronize(S._stream()))


def test_noalloc_rejects_oversize(S):
    V, F = W.icosphere(2, 0.0)
    sc = S.Scene(V, F).build_bvh()
    n = S.lib().snch_closest_silhouette_noalloc  # argument checks happen before any launch
    rc = n(sc._h, ctypes.c_void_p(16), (1 << 22) + 1, 0, None, None, ctypes.c_void_p(16), S._stream())
    assert rc == 1
    assert b"SNCH_NOALLOC_MAX_QUERIES" in S.lib().snch_last_error()
    assert n(None, None, 0, 0, None, None, None, None) == 1
