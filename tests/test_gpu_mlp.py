"""GPU checks of the learned distance bound (csrc/mlp.hip) against the FP64
MLP oracle (oracle/mlp_oracle.py), and of what it gains for the exact query."""
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


def _data(n, seed=0):
    V, F = W.cut_torus(64, 32)
    lo, hi = W.padded_box(V)
    pts = W.uniform_points(n, lo, hi, seed)
    d = np.random.default_rng(seed).uniform(0.0, 0.6, n).astype(np.float32)
    return lo.astype(np.float32), hi.astype(np.float32), pts, d


@pytest.mark.parametrize("tau", [0.5, 0.85])
def test_gradient_matches_fp64(S, tau):
    from oracle import mlp_oracle as M

    lo, hi, pts, d = _data(4096)
    mlp = S.LearnedBound(lo, hi, sigma=4.0, seed=3).set_loss(tau)
    th, four, frame = mlp.params()
    g, loss = mlp.gradient(pts, d)
    lref, gref = M.loss_and_grad(th, pts, d, four, frame, bf16_weights=True, tau=tau)
    assert abs(loss - lref) <= 2e-2 * lref
    blocks = [(M.OFF_W[l], M.OFF_W[l] + M.H * M.KIN[l]) for l in range(4)] + \
             [(M.OFF_B[l], M.OFF_B[l] + M.H) for l in range(4)] + [(M.OFF_W5, M.OFF_B5 + 1)]
    for a, b in blocks:
        rel = np.linalg.norm(g[a:b] - gref[a:b]) / np.linalg.norm(gref[a:b])
        assert rel < 5e-2, (a, b, rel)


def test_inference_matches_fp64(S):
    from oracle import mlp_oracle as M

    lo, hi, pts, d = _data(3000, seed=1)
    mlp = S.LearnedBound(lo, hi, sigma=4.0, seed=5)
    th, four, frame = mlp.params()
    got = mlp.predict(pts).cpu().numpy().astype(np.float64)
    o, _, _ = M.forward(th, pts, four, frame, bf16_weights=True)
    D = frame[7]
    ref = np.maximum(np.exp(o) - 1e-3, 0) * D
    # compare in the network's log domain (BF16 activations: ~1e-2)
    assert np.max(np.abs(np.log(got / D + 1e-3) - np.log(ref / D + 1e-3))) < 5e-2
    r = mlp.seed_radius(pts, rel=0.25, abs_margin=0.01).cpu().numpy()
    np.testing.assert_allclose(r, (got / D * 1.25 + 0.01) * D, rtol=1e-5, atol=1e-6)


def test_training_and_learned_seed(S):
    """Training lowers the loss; a seed radius (learned or exact) never changes
    a result, and an exact seed prunes the traversal."""
    import torch

    V, F = W.cut_torus(128, 64)
    lo, hi = W.padded_box(V)
    sc = S.Scene(V, F).build_bvh()
    n_train = 1 << 17
    pts = W.uniform_points(n_train, lo, hi, 11)
    _, d = sc.closest_silhouette(pts)
    mlp = S.LearnedBound(lo.astype(np.float32), hi.astype(np.float32), sigma=4.0, seed=1)
    losses = mlp.train(pts, d, batch=2048, epochs=20, lr=3e-3)
    assert losses[-1] < 0.5 * losses[0], losses
    q = W.uniform_points(1 << 15, lo, hi, 12)
    st0, st1, st2 = (torch.zeros(5, dtype=torch.int64, device="cuda") for _ in range(3))
    e0, d0 = sc.closest_silhouette(q, stats=st0)
    seed = mlp.seed_radius(q, rel=0.25, abs_margin=0.005)
    e1, d1 = sc.closest_silhouette(q, seed_radius=seed, stats=st1)
    assert torch.equal(e0, e1) and torch.equal(d0, d1)
    # a seed at the exact answer: one capped pass, never a fallback
    e2, d2 = sc.closest_silhouette(q, seed_radius=d0, stats=st2)
    assert torch.equal(e0, e2) and torch.equal(d0, d2)
    assert st2[3].item() < st0[3].item(), (st0.tolist(), st2.tolist())
