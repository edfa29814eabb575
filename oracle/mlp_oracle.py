"""ORACLE (test infrastructure only): FP64 forward/backward of the learned
distance-bound MLP, for checking the GPU kernels in snch_lbvh_amd/csrc/mlp.hip.

Parameter layout (flat FP32 vector, snch_mlp_num_params() entries):
  W1 128x64 @0, b1 @8192, W2 128x128 @8320, b2 @24704, W3 @24832, b3 @41216,
  W4 @41344, b4 @57728, w5 128 @57856, b5 @57984.
Network: f = [sin(B xhat), cos(B xhat)] with B the fixed 32 x 3 Fourier
matrix the library reports and xhat = (x - c) * ih; z_l = W_l a_{l-1} + b_l;
a_l = silu(z_l); o = w5 . a4 + b5; with t = log(min(d / D, 1) + 1e-3) and
err = o - t, loss = mean(w err^2), w = 2 tau if err < 0 else 2 (1 - tau)
(tau = 0.5: the squared error).
The backward pass is pinned by central finite differences in tests."""
import numpy as np

OFF_W = [0, 8320, 24832, 41344]
OFF_B = [8192, 24704, 41216, 57728]
KIN = [64, 128, 128, 128]
OFF_W5, OFF_B5, NPARAM = 57856, 57984, 57985
H = 128


def bf16_round(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta):
    th = np.asarray(theta, np.float64)
    Ws = [th[OFF_W[l]:OFF_W[l] + H * KIN[l]].reshape(H, KIN[l]) for l in range(4)]
    bs = [th[OFF_B[l]:OFF_B[l] + H] for l in range(4)]
    return Ws, bs, th[OFF_W5:OFF_W5 + H], th[OFF_B5]


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def features(x, fourier, frame):
    c = np.asarray(frame[:3], np.float64)
    ih = np.asarray(frame[3:6], np.float64)
    xh = (np.asarray(x, np.float64) - c) * ih
    arg = xh @ np.asarray(fourier, np.float64).reshape(32, 3).T
    return np.concatenate([np.sin(arg), np.cos(arg)], axis=1)


def target(d, frame):
    invD = float(frame[6])
    return np.log(np.minimum(np.asarray(d, np.float64) * invD, 1.0) + 1e-3)


def forward(theta, x, fourier, frame, bf16_weights=False):
    Ws, bs, w5, b5 = unpack(theta)
    if bf16_weights:
        Ws = [bf16_round(W.astype(np.float32)) for W in Ws]
    a = features(x, fourier, frame)
    cache = [(None, a)]
    for l in range(4):
        z = a @ Ws[l].T + bs[l]
        a = silu(z)
        cache.append((z, a))
    return a @ w5 + b5, cache, (Ws, bs, w5, b5)


def loss_and_grad(theta, x, d, fourier, frame, bf16_weights=False, tau=0.5):
    o, cache, (Ws, bs, w5, b5) = forward(theta, x, fourier, frame, bf16_weights)
    t = target(d, frame)
    n = len(o)
    err = o - t
    w = 2.0 * np.where(err < 0, tau, 1.0 - tau)  # expectile weights (tau = 0.5: squared error)
    loss = np.mean(w * err ** 2)
    g = np.zeros(NPARAM)
    go = 2.0 * w * err / n
    a4 = cache[4][1]
    g[OFF_W5:OFF_W5 + H] = go @ a4
    g[OFF_B5] = go.sum()
    delta = go[:, None] * w5[None, :] * dsilu(cache[4][0])
    for l in range(3, -1, -1):
        a_prev = cache[l][1]
        g[OFF_W[l]:OFF_W[l] + H * KIN[l]] = (delta.T @ a_prev).ravel()
        g[OFF_B[l]:OFF_B[l] + H] = delta.sum(0)
        if l > 0:
            delta = (delta @ Ws[l]) * dsilu(cache[l][0])
    return loss, g


def predict(theta, x, fourier, frame):
    o, _, _ = forward(theta, x, fourier, frame)
    D = float(frame[7])
    return np.maximum(np.exp(o) - 1e-3, 0.0) * D
