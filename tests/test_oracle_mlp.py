"""Pins the FP64 numpy network restatement (oracle/mlp.py): its parameter
gradient against central finite differences of its loss, its potential's
position gradient (the orbit force) against finite differences of the
potential, the monopole limit against the closed form, and the target
scaling."""
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


def test_potential_gradient_matches_finite_differences():
    """The orbit force's potential: grad Phi_net from the reverse pass equals
    central finite differences of Phi_net itself (FP64), at radii across the
    training range, with a non-trivial monopole scale."""
    theta = _params(3)
    rng = np.random.default_rng(4)
    u = rng.normal(size=(3, 40))
    q = u / np.linalg.norm(u, axis=0) * np.exp(rng.uniform(0.0, np.log(200.0), 40))
    args = (2.65, 0.65, 0.045, 20.0, 0.3)
    phi, grad = OM.potential_and_gradient(theta, q, *args)
    assert phi.shape == (40,) and grad.shape == (3, 40)
    r = np.linalg.norm(q, axis=0)
    for k in range(3):
        h = 1e-5 * r
        qp, qm = q.copy(), q.copy()
        qp[k] += h
        qm[k] -= h
        fd = (OM.potential_and_gradient(theta, qp, *args)[0] -
              OM.potential_and_gradient(theta, qm, *args)[0]) / (2 * h)
        np.testing.assert_allclose(grad[k], fd, rtol=1e-6, atol=1e-9 * np.abs(grad).max())


def test_potential_is_the_monopole_when_o0_is_zero():
    """With the output layer zeroed, Phi_net is the Hernquist monopole
    -gm / (r + rs) and its gradient the monopole's closed form."""
    p = OM.unpack(_params(5))
    p["W5"][:] = 0.0
    p["b5"][:] = 0.0
    theta = OM.pack(p)
    rng = np.random.default_rng(6)
    q = rng.uniform(-80, 80, (3, 25))
    gm, rs = 0.045, 20.0
    phi, grad = OM.potential_and_gradient(theta, q, 2.65, 0.65, gm, rs, 0.3)
    r = np.linalg.norm(q, axis=0)
    np.testing.assert_allclose(phi, -gm / (r + rs), rtol=1e-14)
    np.testing.assert_allclose(grad, gm / (r + rs) ** 2 * q / r, rtol=1e-13)


def test_targets_invert():
    """targets built from (phi, grad) map back to (phi, grad): o0 through
    the monopole form, o1..o3 through s_grad."""
    rng = np.random.default_rng(2)
    q = rng.uniform(-50, 50, (3, 100))
    phi = -rng.uniform(0.1, 1.0, 100)
    grad = rng.normal(size=(3, 100)) * 0.01
    scale = rng.uniform(0.5, 2.0, 32)
    lnr0, inv_dlnr = 0.0, 31.0 / np.log(200.0)
    gm, rs, c = 2.0, 20.0, 0.25
    feat, target = OM.features_and_targets(q, phi, grad, 2.65, 0.65, lnr0, inv_dlnr, scale,
                                           gm, rs, c)
    r = np.linalg.norm(q, axis=0)
    np.testing.assert_allclose(-gm / (r + rs) * (1.0 + c * target[:, 0]), phi, rtol=1e-12)
    sg = OM.out_scale_grad(r, lnr0, inv_dlnr, scale)
    np.testing.assert_allclose(target[:, 1:].T * sg, grad, rtol=1e-10, atol=1e-14)
    assert feat.shape == (100, 4)


def test_out_scale_interpolates_and_clamps():
    """knot values are reproduced at the knots, midpoints average their
    neighbours, and radii outside the knot span take the end values."""
    scale = np.arange(1.0, 33.0) ** 2
    lnr0, inv_dlnr = np.log(2.0), 2.0
    knots_r = np.exp(lnr0 + np.arange(32) / inv_dlnr)
    np.testing.assert_allclose(OM.out_scale_grad(knots_r, lnr0, inv_dlnr, scale), scale,
                               rtol=1e-12)
    mid = np.exp(lnr0 + (np.arange(31) + 0.5) / inv_dlnr)
    np.testing.assert_allclose(OM.out_scale_grad(mid, lnr0, inv_dlnr, scale),
                               0.5 * (scale[:-1] + scale[1:]), rtol=1e-12)
    sg = OM.out_scale_grad(np.array([1e-3, 1e9]), lnr0, inv_dlnr, scale)
    assert sg[0] == scale[0] and sg[1] == scale[-1]
