"""Walk on stars driven by the library's four queries."""
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


def test_dirichlet_harmonic_kat(S):
    """g = x y is harmonic, so the all-Dirichlet estimator's mean is x0.x * x0.y."""
    V, F = W.cut_torus(256, 128, amp=0.0)
    sc = S.Scene(V, F).build_bvh()
    starts = np.float32([[1.1, 0.2, 0.05], [-0.7, 0.75, -0.1], [0.3, -1.0, 0.0]])
    k = 1 << 14
    x0 = np.repeat(starts, k, axis=0)
    ws = S.WoSt(sc, None, max_walks=len(x0))
    vals, live = ws.solve(x0, max_steps=64, eps=1e-3, seed=7)
    assert live == 0
    v = vals.cpu().numpy().reshape(3, k)
    for s, row in zip(starts, v):
        exact = float(s[0]) * float(s[1])
        se = row.std() / np.sqrt(k)
        assert abs(row.mean() - exact) < 4 * se + 2e-3, (s, row.mean(), exact, se)


def test_mixed_graph_matches_eager(S):
    V, F = W.cut_torus(256, 128)
    FD, FN = W.split_dirichlet_neumann(V, F)
    sd = S.Scene(V, FD).build_bvh()
    sn = S.Scene(V, FN).build_bvh()
    x0 = W.torus_interior_points(1 << 15, seed=8)
    ws = S.WoSt(sd, sn, max_walks=len(x0))
    a, la = ws.solve(x0, max_steps=64, eps=1e-3, seed=3, use_graph=False)
    b, lb = ws.solve(x0, max_steps=64, eps=1e-3, seed=3, use_graph=True)
    c, lc = ws.solve(x0, max_steps=64, eps=1e-3, seed=3, use_graph=True)  # replay
    assert la == lb == lc == 0
    a, b, c = (t.cpu().numpy() for t in (a, b, c))
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(b, c)
    assert np.all(np.isfinite(a))
    # estimates are values of g on the Dirichlet part (z >= 0 there, |x y| <= ~0.9)
    assert np.all(np.abs(a) < 1.0)


def test_mixed_matches_oracle(S, oracle):
    """The CPU restatement walks the same random streams with the oracle's
    queries (breaking distance ties by smallest index, as the library does):
    every walk agrees bit for bit."""
    V, F = W.cut_torus(64, 32)
    FD, FN = W.split_dirichlet_neumann(V, F)
    sd = S.Scene(V, FD).build_bvh()
    sn = S.Scene(V, FN).build_bvh()
    x0 = W.torus_interior_points(4096, seed=8)
    ws = S.WoSt(sd, sn, max_walks=len(x0))
    g, live = ws.solve(x0, max_steps=64, eps=1e-3, seed=3)
    assert live == 0
    o, steps = oracle.wost(oracle.build_tree(V, FD), oracle.build_tree(V, FN), x0, 64, 1e-3, 3)
    g = g.cpu().numpy()
    np.testing.assert_array_equal(g, o)


def test_mixed_slab_kat(S, oracle):
    """Dirichlet g = x y on the sides of the slab [-1,1]^2 x [-h,h], zero
    Neumann flux on its caps: u = x y exactly.  The GPU estimator's mean
    matches x0.x * x0.y at each start, and its walks equal the CPU
    restatement's bit for bit."""
    V, Fs, Fc = W.slab_box(12, 0.25)
    sd = S.Scene(V, Fs).build_bvh()
    sn = S.Scene(V, Fc).build_bvh()
    starts = np.float32([[0.3, -0.5, 0.1], [-0.6, 0.4, -0.2], [0.05, 0.7, 0.22]])
    k = 1 << 15
    x0 = np.repeat(starts, k, axis=0)
    ws = S.WoSt(sd, sn, max_walks=len(x0))
    vals, live = ws.solve(x0, max_steps=256, eps=1e-3, seed=5, use_graph=True)
    assert live == 0
    v = vals.cpu().numpy()
    for s_, row in zip(starts, v.reshape(3, k)):
        exact = float(s_[0]) * float(s_[1])
        se = row.std() / np.sqrt(k)
        assert abs(row.mean() - exact) < 4 * se + 2e-3, (s_, row.mean(), exact, se)
    sub = np.concatenate([np.arange(j * k, j * k + 1024) for j in range(3)])
    o, _ = oracle.wost(oracle.build_tree(V, Fs), oracle.build_tree(V, Fc), x0[sub], 256, 1e-3, 5)
    # the oracle numbers walks from 0: walk j of the subsample is walk sub[j] here
    ws2 = S.WoSt(sd, sn, max_walks=len(sub))
    g2, _ = ws2.solve(x0[sub], max_steps=256, eps=1e-3, seed=5, use_graph=False)
    np.testing.assert_array_equal(g2.cpu().numpy(), o)


def test_graph_recaptured_after_rebuild(S):
    """A captured step loop holds the scenes' tree views; rebuilding the
    Neumann scene with the reference cone radius (which turns the trig-free
    cone test off) must re-capture, so graph and eager results still agree."""
    V, F = W.cut_torus(128, 64)
    FD, FN = W.split_dirichlet_neumann(V, F)
    sd = S.Scene(V, FD).build_bvh()
    sn = S.Scene(V, FN).build_bvh()
    x0 = W.torus_interior_points(1 << 13, seed=8)
    ws = S.WoSt(sd, sn, max_walks=len(x0))
    ws.solve(x0, max_steps=64, eps=1e-3, seed=3, use_graph=True)
    sn.build_bvh(reference_radius=True)
    a, _ = ws.solve(x0, max_steps=64, eps=1e-3, seed=3, use_graph=False)
    b, _ = ws.solve(x0, max_steps=64, eps=1e-3, seed=3, use_graph=True)
    np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
