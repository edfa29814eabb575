"""Pins the FP64 MLP oracle's backward pass with central finite differences."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def _setup(seed=0, n=64):
    rng = np.random.default_rng(seed)
    theta = np.zeros(M.NPARAM)
    for l in range(4):
        theta[M.OFF_W[l]:M.OFF_W[l] + M.H * M.KIN[l]] = rng.uniform(-1, 1, M.H * M.KIN[l]) * np.sqrt(6 / M.KIN[l])
        theta[M.OFF_B[l]:M.OFF_B[l] + M.H] = rng.normal(0, 0.1, M.H)
    theta[M.OFF_W5:M.OFF_W5 + M.H] = rng.uniform(-1, 1, M.H) * np.sqrt(3 / M.H)
    theta[M.OFF_B5] = -2.0
    fourier = rng.normal(0, 4 * 2 * np.pi, 96)
    frame = np.array([0, 0, 0, 1 / 1.3, 1 / 1.3, 1 / 0.4, 1 / 3.0, 3.0])
    x = rng.uniform(-1.3, 1.3, (n, 3)) * np.array([1, 1, 0.3])
    d = rng.uniform(0.001, 0.6, n)
    return theta, x, d, fourier, frame


@pytest.mark.parametrize("tau", [0.5, 0.9])
def test_backward_matches_finite_differences(tau):
    theta, x, d, fourier, frame = _setup()
    _, g = M.loss_and_grad(theta, x, d, fourier, frame, tau=tau)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.integers(M.OFF_W[l], M.OFF_W[l] + M.H * M.KIN[l], 6) for l in range(4)]
                         + [rng.integers(M.OFF_B[l], M.OFF_B[l] + M.H, 3) for l in range(4)]
                         + [rng.integers(M.OFF_W5, M.OFF_W5 + M.H, 4), [M.OFF_B5]])
    h = 1e-6
    for i in idx:
        tp = theta.copy()
        tm = theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (M.loss_and_grad(tp, x, d, fourier, frame, tau=tau)[0]
              - M.loss_and_grad(tm, x, d, fourier, frame, tau=tau)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    x = np.float32([1.0, 1.005859375, 1.00390625, -3.1415927, 1e-20])
    r = M.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0078125 and r[2] == 1.0  # exact halfway rounds to even
    assert abs(r[3] + np.pi) < 2e-2 and r[4] > 0


def test_expectile_weights_closed_form():
    """tau = 0.5 is the plain mean squared error; tau weights the two signs."""
    theta, x, d, fourier, frame = _setup()
    o, _, _ = M.forward(theta, x, fourier, frame)
    t = M.target(d, frame)
    err = o - t
    assert abs(M.loss_and_grad(theta, x, d, fourier, frame, tau=0.5)[0] - np.mean(err ** 2)) < 1e-12
    l9 = M.loss_and_grad(theta, x, d, fourier, frame, tau=0.9)[0]
    # under-prediction (err < 0) weighs 2 tau: the minimiser is an upper expectile
    assert abs(l9 - np.mean(np.where(err < 0, 1.8, 0.2) * err ** 2)) < 1e-12
