"""FP64 numpy restatement of the surrogate network (TEST INFRASTRUCTURE ONLY).

The network has no gala counterpart; this restatement is the checker of the
GPU forward/backward kernels and is itself pinned by central finite
differences of its loss (tests/test_oracle_mlp.py).

Parameters: flat vector, layout W1[4][256] b1[256] W2[256][256] b2 W3 b3 W4
b4 W5[256][4] b5[4] with W[in][out].
Forward:  h1 = silu(f W1 + b1); h_l = silu(h_{l-1} W_l + b_l), l = 2..4;
          o = h4 W5 + b5.
Loss:     L = sum_m sum_o (o_mo - t_mo)^2 / M.
Physics:  Phi_net = Phi_b(r) (1 + c o0) with Phi_b = -gm / (r + rs); the
          orbit force is -grad Phi_net (potential_and_gradient, reverse mode
          in FP64); o1..o3 are auxiliary targets grad Phi / s_grad(r) with the
          radial scale profile interpolated linearly in ln r (out_scale_grad).
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


def out_scale_grad(r, lnr0, inv_dlnr, scale):
    """s_grad(r): knots scale (K,) at ln r = lnr0 + k / inv_dlnr, linear in
    ln r between knots, constant beyond the ends."""
    knots = lnr0 + np.arange(scale.shape[0]) / inv_dlnr
    return np.interp(np.log(r), knots, scale)


def features(q, mu, inv_sigma):
    r = np.sqrt(np.sum(q * q, axis=0))
    return np.stack([q[0] / r, q[1] / r, q[2] / r, (np.log(r) - mu) * inv_sigma], axis=1), r


def features_and_targets(q, phi, grad, mu, inv_sigma, lnr0, inv_dlnr, scale, gm, rs, phi_c):
    """The training pairs: q (3,M), exact phi (M,), grad (3,M).  Features
    (x/r, y/r, z/r, (ln r - mu) inv_sigma); targets ((phi / Phi_b - 1) / c,
    grad / s_grad)."""
    feat, r = features(q, mu, inv_sigma)
    sg = out_scale_grad(r, lnr0, inv_dlnr, scale)
    pb = -gm / (r + rs)
    target = np.stack([(phi / pb - 1.0) / phi_c, grad[0] / sg, grad[1] / sg, grad[2] / sg],
                      axis=1)
    return feat, target


def potential_and_gradient(theta, q, mu, inv_sigma, gm, rs, phi_c):
    """Phi_net (M,) and grad Phi_net (3, M) at q (3, M) in FP64: forward,
    then the reverse pass for d o0 / d f, then the chain rule through the
    features and the monopole."""
    p = unpack(theta)
    feat, r = features(q, mu, inv_sigma)
    y, (f, zs, hs) = forward(theta, feat)
    o0 = y[:, 0]
    d = np.broadcast_to(p["W5"][:, 0], hs[3].shape)        # d o0 / d h4
    for l in range(4, 0, -1):
        s = sigmoid(zs[l - 1])
        dz = d * s * (1.0 + zs[l - 1] * (1.0 - s))           # d o0 / d z_l
        d = dz @ p[f"W{l}"].T                               # d o0 / d (layer l's input)
    g = d                                                   # (M, 4): d o0 / d f
    x, yq, z = q
    gq = (g[:, 0] * x + g[:, 1] * yq + g[:, 2] * z) / r ** 2
    gr = g[:, 3] * inv_sigma / r ** 2
    do = np.stack([(g[:, 0] - gq * x) / r + gr * x, (g[:, 1] - gq * yq) / r + gr * yq,
                   (g[:, 2] - gq * z) / r + gr * z])
    pb = -gm / (r + rs)
    dpb = gm / (r + rs) ** 2
    ratio = 1.0 + phi_c * o0
    return pb * ratio, dpb * ratio * q / r + pb * phi_c * do
