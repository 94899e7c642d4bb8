"""Learned SCF surrogate (no gala counterpart): a 4 -> 256 x 4 (SiLU) -> 4
MLP trained on the exact SCF expansion, on BF16 matrix cores with FP32
accumulation and FP32 master weights, used in a leapfrog whose per-orbit
energy monitor falls back to the exact SCF force.

Everything runs in the HIP kernels of libgala_amd.so (gamd_mlp_*,
gamd_surrogate_leapfrog_step); this module owns the device buffers and the
training loop.

The network's potential is Phi_net = Phi_b(r) (1 + c o0) with Phi_b =
-G M / (r + r_s) the expansion's Hernquist monopole and c the RMS of the
exact Phi / Phi_b - 1 over the training radii, so o0 is of order one; the
orbit force is -grad Phi_net (reverse mode through the network), so
surrogate orbits conserve Phi_net + |v|^2/2 up to the leapfrog's error.
Outputs o1..o3 are auxiliary targets, grad Phi / s_grad(r) with s_grad the
RMS of one exact gradient component over shells in ln r (32 knots between
rmin and rmax, hat-function weights).
"""
import ctypes as C
import math
import time

import numpy as np

from . import _device as dev
from . import _lib


MLPNorm = _lib.MLPNorm

HIDDEN = 256
PARAM_SHAPES = [("W1", (4, HIDDEN)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
                ("W3", (HIDDEN, HIDDEN)), ("b3", (HIDDEN,)), ("W4", (HIDDEN, HIDDEN)),
                ("b4", (HIDDEN,)), ("W5", (HIDDEN, 4)), ("b5", (4,))]
N_PARAM = sum(int(np.prod(s)) for _, s in PARAM_SHAPES)  # 199,684


def init_params(seed=0):
    """FP32 parameters, flat: W ~ N(0, 1/fan_in) (unit-variance inputs to
    SiLU), zero biases, a small output layer."""
    rng = np.random.default_rng(seed)
    out = []
    for name, shape in PARAM_SHAPES:
        if name.startswith("b"):
            out.append(np.zeros(shape))
        else:
            std = 1.0 / math.sqrt(shape[0]) * (0.1 if name == "W5" else 1.0)
            out.append(rng.normal(0.0, std, shape))
    return np.concatenate([a.ravel() for a in out]).astype(np.float32)


def unpack(theta):
    """flat parameters -> dict of arrays (W[in][out])."""
    d, o = {}, 0
    for name, shape in PARAM_SHAPES:
        n = int(np.prod(shape))
        d[name] = np.asarray(theta[o:o + n]).reshape(shape)
        o += n
    return d


class SCFSurrogate:
    """The network for one SCF potential (gala_amd.potential.SCFPotential)."""

    def __init__(self, scf_potential, seed=0, rmin=1.0, rmax=200.0):
        t = dev.require_gpu()
        lib = _lib.load()
        self.lib = lib
        if lib.gamd_mlp_param_count() != N_PARAM:
            raise RuntimeError("parameter layout mismatch with libgala_amd.so")
        self.pot = scf_potential
        self.rmin, self.rmax = rmin, rmax
        lo, hi = math.log(rmin), math.log(rmax)
        self.gm = float(scf_potential.G * scf_potential.parameters["m"])
        self.rs = float(scf_potential.parameters["r_s"])
        self.scale, self.phi_c = self._scale_profile(lo, hi)
        self.norm = MLPNorm(0.5 * (lo + hi), math.sqrt(12.0) / (hi - lo), lo,
                            (_lib.MLP_KNOTS - 1) / (hi - lo), 10.0, self.scale.data_ptr(),
                            self.gm, self.rs, self.phi_c)
        self.theta = dev.to_dev(init_params(seed), dtype=t.float32)
        self.wbf = t.empty(lib.gamd_mlp_bf16_count(), dtype=t.bfloat16, device="cuda")
        self.m = t.zeros_like(self.theta)
        self.v = t.zeros_like(self.theta)
        self.grad = t.zeros_like(self.theta)
        self.step = 0
        self.history = []
        self._sync()

    # ------------------------------------------------------------------
    def _exact(self, q):
        """exact Phi (M,) and grad (3, M) on device at q (3, M) (device or host)."""
        qd = dev.to_dev(q)
        M = qd.shape[1]
        phi = dev.empty((M,))
        grad = dev.empty((3, M))
        s = dev.stream_ptr()
        _lib.check(self.lib.gamd_potential_value(self.pot.c_handle, 0.0, qd.data_ptr(), M,
                                                 phi.data_ptr(), s), "gamd_potential_value")
        _lib.check(self.lib.gamd_potential_gradient(self.pot.c_handle, 0.0, qd.data_ptr(), M,
                                                    grad.data_ptr(), s),
                   "gamd_potential_gradient")
        return qd, phi, grad

    def _scale_profile(self, lo, hi, n=1 << 16, seed=20240):
        """(K,) device table of the RMS of one exact gradient component around
        each of the K knots in ln r (hat-function weights), and c, the RMS of
        Phi / Phi_b - 1 over the sample."""
        K = _lib.MLP_KNOTS
        q = self.sample_points(n, seed)
        _, phi, grad = self._exact(q)
        phi, grad = dev.to_host(phi), dev.to_host(grad)
        r = np.linalg.norm(q, axis=0)
        u = (np.log(r) - lo) * ((K - 1) / (hi - lo))
        wts = np.maximum(0.0, 1.0 - np.abs(u[:, None] - np.arange(K)[None, :]))
        den = np.maximum(wts.sum(axis=0), 1e-300)
        sg = np.sqrt((wts * (np.sum(grad ** 2, axis=0) / 3.0)[:, None]).sum(axis=0) / den)
        if not np.all(np.isfinite(sg)) or np.any(sg <= 0.0):
            raise RuntimeError("degenerate force scale profile (zero force over a shell)")
        c = float(np.sqrt(np.mean((phi / (-self.gm / (r + self.rs)) - 1.0) ** 2)))
        if not (np.isfinite(c) and c > 0.0):
            c = 1.0
        return dev.to_dev(sg), c

    def _sync(self):
        _lib.check(self.lib.gamd_mlp_sync_bf16(self.theta.data_ptr(), self.wbf.data_ptr(),
                                               dev.stream_ptr()), "gamd_mlp_sync_bf16")

    def set_params(self, theta):
        t = dev.torch()
        self.theta = dev.to_dev(np.asarray(theta, dtype=np.float32), dtype=t.float32)
        self._sync()

    def params(self):
        return dev.to_host(self.theta)

    def sample_points(self, n, seed):
        """r log-uniform in [rmin, rmax], isotropic directions; (3, n) FP64."""
        rng = np.random.default_rng(seed)
        r = np.exp(rng.uniform(math.log(self.rmin), math.log(self.rmax), n))
        u = rng.normal(size=(3, n))
        return u / np.linalg.norm(u, axis=0) * r

    def dataset(self, q):
        """features (M,4) and targets (M,4) on device from the exact FP64 SCF."""
        t = dev.torch()
        qd, phi, grad = self._exact(q)
        M = qd.shape[1]
        s = dev.stream_ptr()
        feat = t.empty((M, 4), dtype=t.float32, device="cuda")
        target = t.empty((M, 4), dtype=t.float32, device="cuda")
        _lib.check(self.lib.gamd_mlp_make_targets(C.byref(self.norm), qd.data_ptr(),
                                                  phi.data_ptr(), grad.data_ptr(), M,
                                                  feat.data_ptr(), target.data_ptr(), s),
                   "gamd_mlp_make_targets")
        return feat, target

    def gradients(self, feat, target):
        """(loss, grad) of the batch (M a multiple of 64); grad in self.grad."""
        t = dev.torch()
        M = feat.shape[0]
        ws_bytes = int(self.lib.gamd_mlp_workspace_bytes(M))
        if getattr(self, "_ws", None) is None or self._ws.numel() < ws_bytes:
            self._ws = t.empty(ws_bytes, dtype=t.uint8, device="cuda")
            self._loss = t.zeros(1, dtype=t.float32, device="cuda")
        _lib.check(self.lib.gamd_mlp_gradients(self.theta.data_ptr(), self.wbf.data_ptr(),
                                               feat.data_ptr(), target.data_ptr(), M,
                                               self.grad.data_ptr(), self._loss.data_ptr(),
                                               self._ws.data_ptr(), self._ws.numel(),
                                               dev.stream_ptr()), "gamd_mlp_gradients")
        return self._loss, self.grad

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        _lib.check(self.lib.gamd_mlp_adam(self.theta.data_ptr(), self.grad.data_ptr(),
                                          self.m.data_ptr(), self.v.data_ptr(), self.step, lr,
                                          beta1, beta2, eps, self.wbf.data_ptr(),
                                          dev.stream_ptr()), "gamd_mlp_adam")

    def train(self, n_points=16 * 1024 * 1024, batch=32768, epochs=4, lr=2e-3, lr_min=2e-5,
              seed=1, log=None):
        """Adam with cosine decay over n_points exact-SCF samples."""
        t = dev.torch()
        q = self.sample_points(n_points, seed)
        feat, target = self.dataset(q)
        del q
        n_batches = n_points // batch
        total = epochs * n_batches
        gen = t.Generator(device="cuda")
        gen.manual_seed(seed)
        t0 = time.perf_counter()
        for ep in range(epochs):
            perm = t.randperm(n_points, device="cuda", generator=gen)
            fe, ta = feat[perm].contiguous(), target[perm].contiguous()
            acc = t.zeros(1, dtype=t.float32, device="cuda")
            for b in range(n_batches):
                k = ep * n_batches + b
                cur = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * k / total))
                loss, _ = self.gradients(fe[b * batch:(b + 1) * batch],
                                         ta[b * batch:(b + 1) * batch])
                acc += loss
                self.adam(cur)
            mse = float(acc.item()) / (n_batches * batch * 4)
            self.history.append(mse)
            if log:
                log(f"epoch {ep + 1}/{epochs}: mse {mse:.3e} ({time.perf_counter() - t0:.1f} s)")
        return self.history

    # ------------------------------------------------------------------
    def forward(self, q):
        """Phi_net (M,) and grad Phi_net (3, M) at q (3, M): the network's
        potential and its exact gradient (the negative orbit force)."""
        qd = dev.to_dev(q)
        M = qd.shape[1]
        phi = dev.empty((M,))
        grad = dev.empty((3, M))
        _lib.check(self.lib.gamd_mlp_forward(self.theta.data_ptr(), self.wbf.data_ptr(),
                                             C.byref(self.norm), qd.data_ptr(), M,
                                             phi.data_ptr(), grad.data_ptr(), dev.stream_ptr()),
                   "gamd_mlp_forward")
        return phi, grad

    def leapfrog(self, w0, dt, n_steps, tol=1e-5, progress=None, check_every=100,
                 record_switch=False):
        """Leapfrog of orbits w0 (6, N) with the surrogate force -grad Phi_net
        and per-orbit exact-SCF fallback once |dE/E| of the surrogate energy
        Phi_net + |v|^2/2 exceeds tol (tol < 0: every orbit on the exact
        force from the start).  From its switch step on an orbit follows the
        exact leapfrog bit for bit; every check_every steps, once every orbit
        has switched, the remaining steps run in the exact leapfrog kernel
        resumed from the carried state (gamd_leapfrog_resume).
        Returns dict(w_final (6,N) device, fallback_step (N,) device int32
        [, w_switch (6,N): each switched orbit's state before its switch
        step, position and half-step velocity])."""
        t = dev.torch()
        w = dev.to_dev(w0).clone()
        N = w.shape[1]
        vsync = dev.empty((3, N))
        E0 = t.zeros(N, dtype=t.float64, device="cuda")
        fb = t.full((N,), 0 if tol < 0 else -1, dtype=t.int32, device="cuda")
        wsw = t.full((6, N), float("nan"), dtype=t.float64, device="cuda") if record_switch \
            else None
        s = dev.stream_ptr()
        fn = self.lib.gamd_surrogate_leapfrog_step
        args = (self.theta.data_ptr(), self.wbf.data_ptr(), C.byref(self.norm),
                self.pot.c_handle, w.data_ptr(), vsync.data_ptr(), E0.data_ptr(), fb.data_ptr(),
                dev.ptr(wsw), N, float(dt))
        _lib.check(fn(*args, 0, 0, float(tol), s), "surrogate init")
        out = None
        for j in range(1, n_steps + 1):
            rc = fn(*args, 1, j, float(tol), s)
            if rc:
                _lib.check(rc, "surrogate step")
            if progress and j % 1000 == 0:
                progress(j)
            if j < n_steps and check_every and j % check_every == 0 and \
                    not bool((fb < 0).any().item()):
                out = dev.empty((6, N))
                _lib.check(self.lib.gamd_leapfrog_resume(self.pot.c_handle, w.data_ptr(), N,
                                                         float(dt), n_steps - j, None,
                                                         out.data_ptr(), s),
                           "gamd_leapfrog_resume")
                break
        if out is None:
            out = t.cat([w[:3], vsync], dim=0) if n_steps > 0 else w
        res = {"w_final": out, "fallback_step": fb}
        if record_switch:
            res["w_switch"] = wsw
        return res
