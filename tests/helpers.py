"""Shared builders for the tests: the same physical potentials built once
through the product's classes and once, independently, through the oracle's
raw-parameter helpers."""
import numpy as np

from oracle import oracle as O


def mw_oracle():
    return O.milky_way_like()


def mw_product():
    from gala_amd.potential import MilkyWayLikePotential
    return MilkyWayLikePotential()


def rotation(seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


def shifted_pair(seed=5):
    """A composite with a shifted+rotated Plummer and a rotated MN disk, in
    both descriptions."""
    from gala_amd.potential import (CompositePotential, HernquistPotential,
                                    MiyamotoNagaiPotential, PlummerPotential)
    R1, R2 = rotation(seed), rotation(seed + 1)
    q0 = np.array([1.0, -2.0, 0.5])
    prod = CompositePotential()
    prod["h"] = HernquistPotential(m=3e10, c=2.0)
    prod["p"] = PlummerPotential(m=1e9, b=0.5, origin=q0, R=R1)
    prod["d"] = MiyamotoNagaiPotential(m=5e10, a=2.5, b=0.3, R=R2)
    orc = [O.hernquist(3e10, 2.0), O.plummer(1e9, 0.5, q0=q0, R=R1),
           O.miyamotonagai(5e10, 2.5, 0.3, R=R2)]
    return prod, orc


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def state_rel_err(w, ref):
    """Per-orbit relative error of (6, ..., N) states: position error over
    |x_ref|, velocity error over |v_ref|."""
    dx = np.sqrt(np.sum((w[:3] - ref[:3]) ** 2, axis=0))
    dv = np.sqrt(np.sum((w[3:] - ref[3:]) ** 2, axis=0))
    nx = np.sqrt(np.sum(ref[:3] ** 2, axis=0))
    nv = np.sqrt(np.sum(ref[3:] ** 2, axis=0))
    return np.maximum(dx / nx, dv / nv)


def vcirc_oracle(components):
    def f(q):
        g = O.gradient(components, q)
        r = np.sqrt(np.sum(q ** 2, axis=0))
        return np.sqrt(r * np.abs(np.sum(g * q / r, axis=0)))
    return f
