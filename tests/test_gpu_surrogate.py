"""The learned SCF surrogate on the GPU against the FP64 numpy oracle
(oracle/mlp.py, pinned by finite differences), plus what the trained network
buys over the method without it (the expansion's monopole alone), the
exact-SCF fallback of the surrogate leapfrog, and the full-size surrogate
workload (16M training points, 262,144 orbits x 10,000 steps, 1e-5 bound).

Tolerances: training runs with BF16 weights and activations (8 significant
bits, FP32 accumulation), so its gradients agree with the FP64 restatement
to a few percent per tensor; the orbit force's evaluation splits every
hidden product into three BF16 products (~16 significant bits), so Phi_net
agrees to ~1e-6 of the monopole and its gradient to ~1e-5."""
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


def _net_args(sur):
    nm = sur.norm
    return nm.mu, nm.inv_sigma, nm.gm, nm.rs, nm.phi_c


def test_forward_matches_oracle():
    """Phi_net and grad Phi_net (the negative orbit force) of the GPU's
    precise evaluation (hidden products as three BF16 products of a hi/lo
    split, FP32 accumulation) against the FP64 restatement: a plain BF16
    evaluation is off by ~1e-3 of the monopole part, this one by ~1e-6."""
    from gala_amd import _device as dev
    sur, pot, S, T = _surrogate(3)
    q = sur.sample_points(1000, seed=5)
    phi, grad = sur.forward(q)
    phi, grad = dev.to_host(phi), dev.to_host(grad)
    phi_o, grad_o = OM.potential_and_gradient(sur.params().astype(np.float64), q, *_net_args(sur))
    r = np.linalg.norm(q, axis=0)
    pb = sur.gm / (r + sur.rs)
    ep = np.abs(phi - phi_o) / pb
    eg = np.linalg.norm(grad - grad_o, axis=0) / np.linalg.norm(grad_o, axis=0)
    print(f"Phi_net: max error {ep.max():.2e} of |Phi_b|; grad: median {np.median(eg):.2e}, "
          f"max {eg.max():.2e}")
    assert ep.max() < 1e-5
    assert np.median(eg) < 1e-4 and eg.max() < 1e-3


def test_force_is_the_gradient_of_the_potential():
    """grad Phi_net from the GPU reverse pass equals central finite
    differences of the GPU's own Phi_net: the orbit force is conservative."""
    from gala_amd import _device as dev
    sur, pot, S, T = _surrogate(7)
    q = sur.sample_points(500, seed=8)
    _, grad = sur.forward(q)
    grad = dev.to_host(grad)
    r = np.linalg.norm(q, axis=0)
    h = 1e-4 * r
    fd = np.zeros_like(grad)
    for k in range(3):
        qp, qm = q.copy(), q.copy()
        qp[k] += h
        qm[k] -= h
        fd[k] = (dev.to_host(sur.forward(qp)[0]) - dev.to_host(sur.forward(qm)[0])) / (2 * h)
    e = np.linalg.norm(grad - fd, axis=0) / np.linalg.norm(grad, axis=0)
    print(f"grad vs finite differences of Phi_net: median {np.median(e):.2e}, max {e.max():.2e}")
    assert np.median(e) < 1e-3 and e.max() < 2e-2


def test_make_targets_matches_oracle():
    from gala_amd import _device as dev
    sur, pot, S, T = _surrogate()
    q = sur.sample_points(2000, seed=6)
    feat, target = sur.dataset(q)
    orc = [O.scf(1e12, 20.0, S, T)]
    nm, sc = sur.norm, dev.to_host(sur.scale)
    fo, to = OM.features_and_targets(q, O.value(orc, q), O.gradient(orc, q), nm.mu,
                                     nm.inv_sigma, nm.lnr0, nm.inv_dlnr, sc, nm.gm, nm.rs,
                                     nm.phi_c)
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
    """A short training run: the loss falls, the trained network's force
    -grad Phi_net beats the monopole-only approximation (the method without
    the network), and the surrogate leapfrog's fallback path is the exact
    SCF leapfrog, bit for bit, when forced from the start."""
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
    assert np.median(e_net) < 0.5 * np.median(e_mono)
    # forced fallback (tol < 0) == the exact SCF leapfrog, bit for bit (the
    # handover to the resumed exact kernel after check_every steps included)
    w0, dt, _ = workloads.scf_leapfrog(pot.circular_velocity, n=300, seed=3)
    res = sur.leapfrog(w0, dt, 200, tol=-1.0, check_every=50)
    _, wf = leapfrog(pot, w0, dt, 200, save=False)
    assert np.all(dev.to_host(res["fallback_step"]) == 0)
    assert np.array_equal(dev.to_host(res["w_final"]), dev.to_host(wf))
    # the monitor: a loose bound keeps most orbits on the network, a tight
    # one sends most to the exact force; until its first fallback an orbit
    # evolves identically in both runs, so the tight run never switches later
    fb_loose = dev.to_host(sur.leapfrog(w0, dt, 200, tol=1.0)["fallback_step"])
    fb_tight = dev.to_host(sur.leapfrog(w0, dt, 200, tol=1e-9)["fallback_step"])
    assert np.mean(fb_loose >= 0) < 0.1
    assert np.mean(fb_tight > 0) > 0.5
    both = fb_loose >= 0
    assert np.all((fb_tight[both] >= 0) & (fb_tight[both] <= fb_loose[both]))


def test_config4_full_size():
    """The learned-surrogate workload at full size: 16,777,216 training
    points (8 epochs), then 262,144 orbits x 10,000 steps of 0.5 Myr with the
    1e-5 energy-drift fallback.  Orbits that stay on the network conserve
    their surrogate energy to 1e-5 by construction; every switched orbit
    continues as the exact leapfrog bit for bit from its switch step; the
    true (FP64 SCF) energy errors and the force errors are reported."""
    from gala_amd import _device as dev
    from gala_amd import _lib
    from gala_amd.integrate import leapfrog
    import torch
    sur, pot, S, T = _surrogate(0)
    print("config 4: training on 16,777,216 points", flush=True)
    hist = sur.train(n_points=1 << 24, batch=32768, epochs=8, lr=2e-3, lr_min=2e-5, seed=1)
    assert hist[-1] < 0.05 * hist[0]
    q = sur.sample_points(1 << 20, seed=99)
    _, g_net = sur.forward(q)
    e_f = np.linalg.norm(dev.to_host(g_net) - pot.gradient(q), axis=0) / \
        np.linalg.norm(pot.gradient(q), axis=0)
    w0, dt, n_steps = workloads.scf_leapfrog(pot.circular_velocity, n=262144, seed=3)
    assert n_steps == 10000
    print("config 4: 262,144 orbits x 10,000 steps, tol 1e-5", flush=True)
    res = sur.leapfrog(w0, dt, n_steps, tol=1e-5, record_switch=True)
    fb = dev.to_host(res["fallback_step"])
    wf = res["w_final"]
    frac = np.mean(fb >= 0)
    # true energy errors against the FP64 SCF
    def energy(w):
        lib = _lib.load()
        out = dev.empty((w.shape[1],))
        _lib.check(lib.gamd_hamiltonian_energy(pot.c_handle, 0.0, w.contiguous().data_ptr(),
                                               w.shape[1], out.data_ptr(), dev.stream_ptr()),
                   "energy")
        return dev.to_host(out)
    e0 = energy(dev.to_dev(w0))
    de = np.abs(energy(wf) - e0) / np.abs(e0)
    _, wx = leapfrog(pot, w0, dt, n_steps, save=False)
    de_exact = np.abs(energy(wx) - e0) / np.abs(e0)
    on = fb < 0
    print(f"config 4: {frac:.4f} of orbits fell back (median switch step "
          f"{np.median(fb[~on]) if np.any(~on) else -1:.0f}); force error median "
          f"{np.median(e_f):.2e}, 90% {np.percentile(e_f, 90):.2e}; true |dE/E| of orbits kept "
          f"on the network: median {np.median(de[on]) if np.any(on) else np.nan:.2e}, 90% "
          f"{np.percentile(de[on], 90) if np.any(on) else np.nan:.2e}; all orbits median "
          f"{np.median(de):.2e}; exact leapfrog median {np.median(de_exact):.2e}")
    assert 0.0 < frac < 1.0
    assert np.all(np.isfinite(dev.to_host(wf)))
    # switched orbits: resume the exact kernel from the recorded state before
    # the switch step; the run's final states must be reproduced bit for bit
    sw = np.nonzero(fb >= 1)[0]
    pick = sw[np.linspace(0, len(sw) - 1, min(len(sw), 4096)).astype(int)]
    wsw = res["w_switch"][:, torch.as_tensor(pick, device="cuda")].contiguous()
    nsp = dev.to_dev(n_steps - fb[pick] + 1, dtype=torch.int64)
    out = dev.empty((6, len(pick)))
    lib = _lib.load()
    _lib.check(lib.gamd_leapfrog_resume(pot.c_handle, wsw.data_ptr(), len(pick), float(dt), 0,
                                        nsp.data_ptr(), out.data_ptr(), dev.stream_ptr()),
               "gamd_leapfrog_resume")
    assert np.array_equal(dev.to_host(out), dev.to_host(wf)[:, pick])
