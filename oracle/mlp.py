"""FP64 numpy restatement of the surrogate network (TEST INFRASTRUCTURE ONLY).

The network has no gala counterpart; this restatement is the checker of the
GPU forward/backward kernels and is itself pinned by central finite
differences of its loss (tests/test_oracle_mlp.py).

Parameters: flat vector, layout W1[4][256] b1[256] W2[256][256] b2 W3 b3 W4
b4 W5[256][4] b5[4] with W[in][out].
Forward:  h1 = silu(f W1 + b1); h_l = silu(h_{l-1} W_l + b_l), l = 2..4;
          o = h4 W5 + b5.
Loss:     L = sum_m sum_o (o_mo - t_mo)^2 / M.
Physics:  Phi = s_phi(r) o0, grad = s_grad(r) o[1:4] with the radial scale
          profiles interpolated linearly in ln r (out_scales).
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


def out_scales(r, lnr0, inv_dlnr, scale):
    """radial output scales: scale (2, K) knots at ln r = lnr0 + k / inv_dlnr,
    linear in ln r between knots, constant beyond the ends."""
    knots = lnr0 + np.arange(scale.shape[1]) / inv_dlnr
    lr = np.log(r)
    return np.interp(lr, knots, scale[0]), np.interp(lr, knots, scale[1])


def features_and_targets(q, phi, grad, mu, inv_sigma, lnr0, inv_dlnr, scale):
    """The scaled training pairs: q (3,M), exact phi (M,), grad (3,M).
    Features (x/r, y/r, z/r, (ln r - mu) inv_sigma)."""
    r = np.sqrt(np.sum(q * q, axis=0))
    feat = np.stack([q[0] / r, q[1] / r, q[2] / r, (np.log(r) - mu) * inv_sigma], axis=1)
    sp, sg = out_scales(r, lnr0, inv_dlnr, scale)
    target = np.stack([phi / sp, grad[0] / sg, grad[1] / sg, grad[2] / sg], axis=1)
    return feat, target


def physical(q, y, lnr0, inv_dlnr, scale):
    """network outputs -> (Phi, grad) at q."""
    r = np.sqrt(np.sum(q * q, axis=0))
    sp, sg = out_scales(r, lnr0, inv_dlnr, scale)
    return sp * y[:, 0], sg * y[:, 1:4].T
