"""Seeded synthetic inputs of the benchmark workloads.

All draws come from one ``numpy.random.default_rng(seed)`` in a fixed order
(written out below), in galactic units (kpc, Myr, Msun; velocities kpc/Myr).
Circular speeds come from the potential's own gradient
(circular_velocity, gala potential/potential/core.py:772-785), evaluated by
the caller-supplied function so that the same draws can be made with the GPU
path or with a CPU checker.

Workloads:
  composite_leapfrog   1024 orbits, dt = 1 Myr, 10,000 steps in the
                       Milky-Way-like composite; r ~ U(5, 30) kpc isotropic;
                       v = v_c N(1, 0.1) along a random tangent + N(0, 10 km/s)
  composite_dop853     1,048,576 orbits, rtol 1e-8, atol 1e-10; integration
                       time logU(0.1, 10) Gyr; r ~ logU(1, 200) kpc; speed
                       factor N(1, 0.3)
  scf_leapfrog         262,144 orbits in the SCF potential (nmax 10, lmax 6,
                       M 1e12, r_s 20); dt 0.5 Myr, 10,000 steps; r ~ logU(2, 150)
  mockstream           Plummer progenitor (1e5 Msun, 10 pc) at (15, 0, 0) kpc,
                       (0, 180, 30) km/s, 4000 steps back at dt = -1 Myr
"""
import numpy as np

from .units import KMS_TO_KPCMYR

MW_COMPONENTS = (("hernquist", {"m": 1.7e9, "c": 0.07}),
                 ("hernquist", {"m": 5e9, "c": 1.0}),
                 ("miyamotonagai", {"m": 6.8e10, "a": 3.0, "b": 0.28}),
                 ("nfw", {"m": 5.4e11, "r_s": 15.6}))


def _isotropic(rng, n):
    """unit vectors: cos(theta) ~ U(-1, 1), phi ~ U(0, 2 pi)"""
    cth = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    sth = np.sqrt(1.0 - cth * cth)
    return np.stack([sth * np.cos(phi), sth * np.sin(phi), cth])


def _tangent(rng, rhat):
    """random unit vector perpendicular to each column of rhat: angle psi ~
    U(0, 2 pi) in the plane spanned by e1 = normalised (z_hat x r_hat) (x_hat
    for r_hat on the z axis) and e2 = r_hat x e1."""
    n = rhat.shape[1]
    psi = rng.uniform(0.0, 2.0 * np.pi, n)
    zhat = np.array([0.0, 0.0, 1.0])[:, None]
    e1 = np.cross(zhat.T, rhat.T).T
    norm = np.linalg.norm(e1, axis=0)
    bad = norm < 1e-12
    e1[:, bad] = np.array([1.0, 0.0, 0.0])[:, None]
    norm[bad] = 1.0
    e1 /= norm
    e2 = np.cross(rhat.T, e1.T).T
    return np.cos(psi) * e1 + np.sin(psi) * e2


def initial_conditions(rng, n, r, vcirc_fn, speed_sigma, noise_kms=10.0):
    """Positions at radii r along isotropic directions; velocities
    v_c(q) * N(1, speed_sigma) along a random tangent plus N(0, noise) km/s per
    component.  Draw order: directions, tangent angle, speed factor, noise."""
    rhat = _isotropic(rng, n)
    q = r * rhat
    tang = _tangent(rng, rhat)
    fac = rng.normal(1.0, speed_sigma, n)
    noise = rng.normal(0.0, noise_kms * KMS_TO_KPCMYR, (3, n))
    vc = vcirc_fn(q)
    v = vc * fac * tang + noise
    return np.ascontiguousarray(np.vstack([q, v]))


def composite_leapfrog(vcirc_fn, n=1024, seed=0):
    """(w0 (6,n), dt, n_steps) for the 1024-orbit leapfrog workload."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(5.0, 30.0, n)
    return initial_conditions(rng, n, r, vcirc_fn, 0.1), 1.0, 10000


def composite_dop853(vcirc_fn, n=1048576, seed=1):
    """(w0 (6,n), T (n,)) for the uneven-length DOP853 workload; each orbit
    runs from t = 0 to T_i with T ~ logU(100, 10,000) Myr."""
    rng = np.random.default_rng(seed)
    r = np.exp(rng.uniform(np.log(1.0), np.log(200.0), n))
    w0 = initial_conditions(rng, n, r, vcirc_fn, 0.3)
    T = np.exp(rng.uniform(np.log(100.0), np.log(10000.0), n))
    return w0, T


def scf_coefficients(nmax=10, lmax=6, seed=3):
    """S_nlm = N(0,1) (n+1)^-2 (l+1)^-2 with S_000 = 1; T likewise with
    T_nl0 = 0; entries with m > l zeroed."""
    rng = np.random.default_rng(seed)
    n = np.arange(nmax + 1)[:, None, None]
    l = np.arange(lmax + 1)[None, :, None]
    m = np.arange(lmax + 1)[None, None, :]
    scale = 1.0 / ((n + 1.0) ** 2 * (l + 1.0) ** 2)
    S = rng.normal(0.0, 1.0, (nmax + 1, lmax + 1, lmax + 1)) * scale
    T = rng.normal(0.0, 1.0, (nmax + 1, lmax + 1, lmax + 1)) * scale
    valid = (m <= l) * np.ones_like(S)
    S *= valid
    T *= valid
    T[:, :, 0] = 0.0
    S[0, 0, 0] = 1.0
    return np.ascontiguousarray(S), np.ascontiguousarray(T)


def scf_leapfrog(vcirc_fn, n=262144, seed=3):
    """(w0 (6,n), dt, n_steps) for the SCF leapfrog workload (radii from the
    same generator after the coefficients' draws use their own seed)."""
    rng = np.random.default_rng(seed + 100)
    r = np.exp(rng.uniform(np.log(2.0), np.log(150.0), n))
    return initial_conditions(rng, n, r, vcirc_fn, 0.1), 0.5, 10000


def mockstream_progenitor():
    """(prog_w0 (6,), prog_mass, plummer b, dt, n_steps)"""
    w0 = np.array([15.0, 0.0, 0.0, 0.0, 180.0 * KMS_TO_KPCMYR, 30.0 * KMS_TO_KPCMYR])
    return w0, 1e5, 0.01, -1.0, 4000


def cost_order(w0, T, gradient_fn, spread=1, per_wave=4, n_spread=2048):
    """Longest-first queue order for the DOP853 work queue: the estimated
    number of dynamical times T_i * sqrt(|g| / r) per orbit, descending.

    The n_spread most expensive orbits are dealt per_wave at a time to the
    first slots of consecutive spread-slot blocks instead of filling the head
    of the queue: the persistent kernel's waves first claim 64 consecutive
    slots each, so each of n_spread / per_wave waves starts with per_wave of
    them among cheaper orbits, holds only those once the queue is drained,
    and finishes them in the kernel's wide tail (up to four orbits on 16-lane
    rows).  spread <= 1 (the default) or n_spread = 0 keeps the plain
    longest-first order, the one to pair with gamd_dop853's n_tail, which
    gives the head of the order dedicated 16-lane rows from the start."""
    q = w0[:3]
    r = np.sqrt(np.sum(q * q, axis=0))
    g = gradient_fn(q)
    om = np.sqrt(np.sqrt(np.sum(g * g, axis=0)) / r)
    cost = np.abs(T) * om
    base = np.argsort(-cost, kind="stable").astype(np.int32)
    N = base.shape[0]
    if spread <= 1 or per_wave < 1 or per_wave > spread:
        return base
    k = min(n_spread, (N // spread) * per_wave)
    k -= k % per_wave
    if k <= 0:
        return base
    blocks = np.arange(k // per_wave) * spread
    slots = (blocks[:, None] + np.arange(per_wave)[None, :]).ravel()
    rest = np.ones(N, dtype=bool)
    rest[slots] = False
    out = np.empty_like(base)
    out[slots] = base[:k]
    out[rest] = base[k:]
    return out