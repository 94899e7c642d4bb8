"""Pins the FP64 numpy network restatement (oracle/mlp.py): its analytic
gradient against central finite differences of its loss, and the output
scaling against the closed-form monopole."""
import numpy as np

from oracle import mlp as OM


def _params(seed):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in OM.SHAPES:
        scale = 1.0 / np.sqrt(shape[0]) if name.startswith("W") else 0.1
        d[name] = rng.normal(0.0, scale, shape)
    return OM.pack(d)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    theta = _params(1)
    feat = rng.normal(size=(16, 4))
    target = rng.normal(size=(16, 4)) * 0.1
    L, g = OM.loss_and_grad(theta, feat, target)
    # spot-check entries of every tensor
    offsets = np.cumsum([0] + [int(np.prod(s)) for _, s in OM.SHAPES])
    h = 1e-6
    for a, b in zip(offsets[:-1], offsets[1:]):
        for idx in rng.choice(np.arange(a, b), size=min(6, b - a), replace=False):
            tp = theta.copy()
            tp[idx] += h
            tm = theta.copy()
            tm[idx] -= h
            fd = (OM.loss_and_grad(tp, feat, target)[0] - OM.loss_and_grad(tm, feat, target)[0]) / (2 * h)
            assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (idx, fd, g[idx])


def test_physical_scaling_inverts_targets():
    """targets built from (phi, grad) map back to (phi, grad) exactly."""
    rng = np.random.default_rng(2)
    q = rng.uniform(-50, 50, (3, 100))
    phi = -rng.uniform(0.1, 1.0, 100)
    grad = rng.normal(size=(3, 100)) * 0.01
    scale = rng.uniform(0.5, 2.0, (2, 32))
    lnr0, inv_dlnr = 0.0, 31.0 / np.log(200.0)
    feat, target = OM.features_and_targets(q, phi, grad, 2.65, 0.65, lnr0, inv_dlnr, scale)
    p2, g2 = OM.physical(q, target, lnr0, inv_dlnr, scale)
    np.testing.assert_allclose(p2, phi, rtol=1e-12)
    np.testing.assert_allclose(g2, grad, rtol=1e-10, atol=1e-14)
    assert feat.shape == (100, 4)


def test_out_scales_interpolate_and_clamp():
    """knot values are reproduced at the knots, midpoints average their
    neighbours, and radii outside the knot span take the end values."""
    scale = np.stack([np.arange(1.0, 33.0), np.arange(1.0, 33.0) ** 2])
    lnr0, inv_dlnr = np.log(2.0), 2.0
    knots_r = np.exp(lnr0 + np.arange(32) / inv_dlnr)
    sp, sg = OM.out_scales(knots_r, lnr0, inv_dlnr, scale)
    np.testing.assert_allclose(sp, scale[0], rtol=1e-12)
    np.testing.assert_allclose(sg, scale[1], rtol=1e-12)
    mid = np.exp(lnr0 + (np.arange(31) + 0.5) / inv_dlnr)
    sp, _ = OM.out_scales(mid, lnr0, inv_dlnr, scale)
    np.testing.assert_allclose(sp, 0.5 * (scale[0][:-1] + scale[0][1:]), rtol=1e-12)
    sp, sg = OM.out_scales(np.array([1e-3, 1e9]), lnr0, inv_dlnr, scale)
    assert sp[0] == scale[0][0] and sp[1] == scale[0][-1]
    assert sg[0] == scale[1][0] and sg[1] == scale[1][-1]
