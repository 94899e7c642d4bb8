"""FP64 numpy restatement of the surrogate network (TEST INFRASTRUCTURE ONLY).

The network has no gala counterpart; this restatement is the checker of the
GPU forward/backward kernels and is itself pinned by central finite
differences of its loss (tests/test_oracle_mlp.py).

Parameters: flat vector, layout W1[4][256] b1[256] W2[256][256] b2 W3 b3 W4
b4 W5[256][4] b5[4] with W[in][out].
Forward:  h1 = silu(f W1 + b1); h_l = silu(h_{l-1} W_l + b_l), l = 2..4;
          o = h4 W5 + b5.
Loss:     L = sum_m sum_o (o_mo - t_mo)^2 / M.
"""
import numpy as np

H = 256
SHAPES = [("W1", (4, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)),
          ("b3", (H,)), ("W4", (H, H)), ("b4", (H,)), ("W5", (H, 4)), ("b5", (4,))]


def unpack(theta):
    d, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        d[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return d


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(theta, feat):
    p = unpack(theta)
    zs, hs = [], []
    h = np.asarray(feat, dtype=np.float64)
    for l in range(1, 5):
        z = h @ p[f"W{l}"] + p[f"b{l}"]
        h = z * sigmoid(z)
        zs.append(z)
        hs.append(h)
    y = h @ p["W5"] + p["b5"]
    return y, (np.asarray(feat, dtype=np.float64), zs, hs)


def loss_and_grad(theta, feat, target):
    p = unpack(theta)
    y, (f, zs, hs) = forward(theta, feat)
    M = y.shape[0]
    d = y - target
    L = np.sum(d * d) / M
    dy = 2.0 * d / M
    g = {"W5": hs[3].T @ dy, "b5": dy.sum(axis=0)}
    dh = dy @ p["W5"].T
    for l in range(4, 0, -1):
        s = sigmoid(zs[l - 1])
        dz = dh * s * (1.0 + zs[l - 1] * (1.0 - s))
        hin = hs[l - 2] if l > 1 else f
        g[f"W{l}"] = hin.T @ dz
        g[f"b{l}"] = dz.sum(axis=0)
        dh = dz @ p[f"W{l}"].T
    return L, pack(g)


def features_and_targets(q, phi, grad, L, mu, inv_sigma, GM, r_s):
    """The scaled training pairs: q (3,M), exact phi (M,), grad (3,M)."""
    r = np.sqrt(np.sum(q * q, axis=0))
    feat = np.stack([q[0] / L, q[1] / L, q[2] / L, (np.log(r) - mu) * inv_sigma], axis=1)
    rs = r + r_s
    phiH = -GM / rs
    gH = GM / rs ** 2
    target = np.stack([phi / phiH - 1.0, grad[0] / gH - q[0] / r, grad[1] / gH - q[1] / r,
                       grad[2] / gH - q[2] / r], axis=1)
    return feat, target


def physical(q, y, GM, r_s):
    """network outputs -> (Phi, grad) at q."""
    r = np.sqrt(np.sum(q * q, axis=0))
    rs = r + r_s
    phi = -GM / rs * (1.0 + y[:, 0])
    gH = GM / rs ** 2
    grad = gH * (q / r + y[:, 1:4].T)
    return phi, grad
