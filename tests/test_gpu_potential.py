"""GPU potential evaluation (gamd_potential_*) against the oracle.

Tolerances: the HIP path contracts multiply-adds into FMAs and uses its own
sqrt/log, so agreement is to a few ulp, checked at 1e-13 relative (1e-12 for
the second difference, whose cancellation amplifies ulp differences by
~1/(h^2 Phi'') relative to Phi)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import mw_oracle, mw_product, rel_err, shifted_pair

pytestmark = pytest.mark.gpu


def _points(seed, n=4096, lo=-40, hi=40):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=(3, n))


@pytest.mark.parametrize("which", ["mw", "shifted"])
def test_value_gradient(which):
    if which == "mw":
        prod, orc = mw_product(), mw_oracle()
    else:
        prod, orc = shifted_pair()
    q = _points(1)
    v = prod.energy(q)
    g = prod.gradient(q)
    vo = O.value(orc, q)
    go = O.gradient(orc, q)
    assert np.max(rel_err(v, vo)) < 1e-13
    gn = np.sqrt(np.sum(go ** 2, axis=0))
    assert np.max(np.abs(g - go) / gn) < 1e-13


def test_density_hessian_d2():
    prod, orc = mw_product(), mw_oracle()
    q = _points(2, n=1024)
    assert np.max(rel_err(prod.density(q), O.density(orc, q))) < 1e-12
    h = prod.hessian(q)
    ho = O.hessian(orc, q)
    scale = np.max(np.abs(ho), axis=(0, 1))
    assert np.max(np.abs(h - ho) / scale) < 1e-12
    d2 = prod.d2_dr2(q)
    d2o = O.d2_dr2(orc, q)
    # (Phi(r+h) - 2 Phi(r) + Phi(r-h)) / h^2 with h = 1e-2 cancels: a few ulp
    # of Phi in each term become 4 * 16 ulp(Phi) / h^2 in the result
    phi = np.abs(O.value(orc, q))
    assert np.all(np.abs(d2 - d2o) <= 64 * np.finfo(float).eps * phi / 1e-4)


def test_energy_and_shapes():
    from gala_amd.potential import Hamiltonian
    prod, orc = mw_product(), mw_oracle()
    rng = np.random.default_rng(3)
    w = rng.uniform(-10, 10, size=(6, 7, 5))
    E = Hamiltonian(prod).energy(w)
    assert E.shape == (7, 5)
    w2 = w.reshape(6, -1)
    Eo = O.value(orc, w2[:3]) + 0.5 * np.sum(w2[3:] ** 2, axis=0)
    assert np.max(rel_err(E.ravel(), Eo)) < 1e-13
    assert prod.gradient(np.array([1.0, 2.0, 3.0])).shape == (3,)
    assert prod.energy(np.zeros((3, 2, 3)) + 1.0).shape == (2, 3)


def test_nfw_special_cases():
    from gala_amd.potential import NFWPotential
    p = NFWPotential(m=5.4e11, r_s=15.6)
    q = np.zeros((3, 1))
    # value at r = 0 is -G m / r_s (builtin_potentials.cpp:828-829); the
    # gradient is 0/0 there (:844-845)
    assert p.energy(q)[0] == -p.G * 5.4e11 / 15.6
    assert np.all(np.isnan(p.gradient(q)))
