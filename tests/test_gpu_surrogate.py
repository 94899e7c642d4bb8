"""The learned SCF-force surrogate on the GPU against the FP64 numpy oracle
(oracle/mlp.py, pinned by finite differences), plus what the trained network
buys over the method without it (the expansion's monopole alone), and the
exact-SCF fallback of the surrogate leapfrog.

Tolerances: weights and activations are BF16 (8 significant bits) with FP32
accumulation, so the network outputs agree with the FP64 restatement to
~1e-2 of their scale and gradients to a few percent in norm."""
import numpy as np
import pytest

from gala_amd import workloads
from oracle import mlp as OM
from oracle import oracle as O
from tests.helpers import state_rel_err

pytestmark = pytest.mark.gpu


def _surrogate(seed=0):
    from gala_amd.potential import SCFPotential
    from gala_amd.surrogate import SCFSurrogate
    S, T = workloads.scf_coefficients()
    pot = SCFPotential(m=1e12, r_s=20.0, Snlm=S, Tnlm=T)
    return SCFSurrogate(pot, seed=seed), pot, S, T


def test_forward_matches_oracle():
    from gala_amd import _device as dev
    sur, pot, S, T = _surrogate(3)
    q = sur.sample_points(1000, seed=5)
    phi, grad = sur.forward(q)
    phi, grad = dev.to_host(phi), dev.to_host(grad)
    nm, sc = sur.norm, dev.to_host(sur.scale)
    feat, _ = OM.features_and_targets(q, np.ones(1000), np.ones((3, 1000)), nm.mu,
                                      nm.inv_sigma, nm.lnr0, nm.inv_dlnr, sc)
    y, _ = OM.forward(sur.params(), feat)
    phi_o, grad_o = OM.physical(q, y, nm.lnr0, nm.inv_dlnr, sc)
    sp, sg = OM.out_scales(np.linalg.norm(q, axis=0), nm.lnr0, nm.inv_dlnr, sc)
    # compare in network-output units
    assert np.max(np.abs(phi - phi_o) / sp) < 2e-2 * (1 + np.max(np.abs(y[:, 0])))
    assert np.max(np.abs(grad - grad_o) / sg) < 2e-2 * (1 + np.max(np.abs(y[:, 1:])))


def test_make_targets_matches_oracle():
    from gala_amd import _device as dev
    sur, pot, S, T = _surrogate()
    q = sur.sample_points(2000, seed=6)
    feat, target = sur.dataset(q)
    orc = [O.scf(1e12, 20.0, S, T)]
    nm, sc = sur.norm, dev.to_host(sur.scale)
    fo, to = OM.features_and_targets(q, O.value(orc, q), O.gradient(orc, q), nm.mu,
                                     nm.inv_sigma, nm.lnr0, nm.inv_dlnr, sc)
    np.testing.assert_allclose(dev.to_host(feat), fo, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(dev.to_host(target), to, rtol=1e-5, atol=1e-6)


def test_gradients_match_oracle():
    from gala_amd import _device as dev
    sur, pot, S, T = _surrogate(4)
    rng = np.random.default_rng(0)
    M = 1024
    feat = rng.normal(size=(M, 4)).astype(np.float32)
    target = (rng.normal(size=(M, 4)) * 0.3).astype(np.float32)
    loss, g = sur.gradients(dev.to_dev(feat, dtype=dev.torch().float32),
                            dev.to_dev(target, dtype=dev.torch().float32))
    g = dev.to_host(g).astype(np.float64)
    L, go = OM.loss_and_grad(sur.params(), feat, target)
    assert abs(float(loss.item()) / M - L) < 2e-2 * L
    cos = g @ go / (np.linalg.norm(g) * np.linalg.norm(go))
    assert cos > 0.995
    offsets = np.cumsum([0] + [int(np.prod(s)) for _, s in OM.SHAPES])
    for (name, _), a, b in zip(OM.SHAPES, offsets[:-1], offsets[1:]):
        rel = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        assert rel < 6e-2, (name, rel)


def test_training_beats_monopole_and_falls_back():
    """A short training run: the loss falls, the trained network's forces
    beat the monopole-only approximation (the method without the network),
    and the surrogate leapfrog's fallback path reproduces the exact SCF
    leapfrog when forced from the start."""
    from gala_amd import _device as dev
    from gala_amd.integrate import leapfrog
    sur, pot, S, T = _surrogate(0)
    hist = sur.train(n_points=1 << 21, batch=8192, epochs=3, lr=2e-3, lr_min=1e-4, seed=1)
    assert hist[-1] < 0.1 * hist[0]
    q = sur.sample_points(20000, seed=99)
    _, g_net = sur.forward(q)
    g_net = dev.to_host(g_net)
    g_ex = pot.gradient(q)
    r = np.linalg.norm(q, axis=0)
    GM, r_s = pot.G * pot.parameters["m"], pot.parameters["r_s"]
    g_mono = GM / (r + r_s) ** 2 * q / r  # the expansion's S_000 (Hernquist) term alone
    e_net = np.linalg.norm(g_net - g_ex, axis=0) / np.linalg.norm(g_ex, axis=0)
    e_mono = np.linalg.norm(g_mono - g_ex, axis=0) / np.linalg.norm(g_ex, axis=0)
    print(f"median force error: network {np.median(e_net):.2e}, monopole {np.median(e_mono):.2e}")
    assert np.median(e_net) < 0.3 * np.median(e_mono)
    # forced fallback (tol < 0) == the exact SCF leapfrog
    w0, dt, _ = workloads.scf_leapfrog(pot.circular_velocity, n=300, seed=3)
    res = sur.leapfrog(w0, dt, 200, tol=-1.0)
    _, wf = leapfrog(pot, w0, dt, 200, save=False)
    assert np.all(dev.to_host(res["fallback_step"]) == 0)
    assert np.max(state_rel_err(dev.to_host(res["w_final"]), dev.to_host(wf))) < 1e-10
    # the monitor: a loose bound lets most orbits stay on the network (those
    # that leave the trained radii fall back), a tight one sends most to the
    # exact force; until its first fallback an orbit evolves identically in
    # both runs, so the tight run never switches later than the loose one
    fb_loose = dev.to_host(sur.leapfrog(w0, dt, 200, tol=1.0)["fallback_step"])
    fb_tight = dev.to_host(sur.leapfrog(w0, dt, 200, tol=1e-9)["fallback_step"])
    assert np.mean(fb_loose >= 0) < 0.1
    assert np.mean(fb_tight > 0) > 0.5
    both = fb_loose >= 0
    assert np.all((fb_tight[both] >= 0) & (fb_tight[both] <= fb_loose[both]))
