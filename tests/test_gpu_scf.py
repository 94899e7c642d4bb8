"""GPU SCF expansion (the SCF component of the potential table) against the
Fortran golden vectors (rtol 1e-6, as gala's test_accp_fortran.py) and the
oracle (1e-12), plus SCF orbits against the oracle."""
import numpy as np
import pytest

from gala_amd import workloads
from oracle import oracle as O
from tests.helpers import rel_err, state_rel_err
from tests.test_oracle_scf import NAMES, fortran_case

pytestmark = pytest.mark.gpu


def product_scf(m, r_s, S, T, G=None):
    from gala_amd.potential import SCFPotential
    from gala_amd.units import UnitSystem
    units = UnitSystem(G=G) if G is not None else None
    return SCFPotential(m=m, r_s=r_s, Snlm=S, Tnlm=T, units=units)


@pytest.mark.parametrize("name", NAMES)
def test_fortran_golden(name):
    xyz, S, T, accp = fortran_case(name)
    pot = product_scf(1.0, 1.0, S, T, G=1.0)
    np.testing.assert_allclose(pot.energy(xyz), -accp[:, -1], rtol=1e-6)
    g = pot.gradient(xyz)
    np.testing.assert_allclose(g.T, accp[:, :3], rtol=1e-6, atol=1e-7 * np.abs(accp[:, :3]).max())


def test_oracle_parity_benchmark_coefficients():
    S, T = workloads.scf_coefficients()
    pot = product_scf(1e12, 20.0, S, T)
    orc = [O.scf(1e12, 20.0, S, T)]
    rng = np.random.default_rng(0)
    r = np.exp(rng.uniform(np.log(1.0), np.log(200.0), 2000))
    q = r * rng.normal(size=(3, 2000)) / np.linalg.norm(rng.normal(size=(3, 2000)), axis=0)
    q = r / np.linalg.norm(q, axis=0) * q
    assert np.max(rel_err(pot.energy(q), O.value(orc, q))) < 1e-12
    g, go = pot.gradient(q), O.gradient(orc, q)
    assert np.max(np.abs(g - go) / np.linalg.norm(go, axis=0)) < 1e-11
    rho, rho_o = pot.density(q), O.density(orc, q)
    assert np.max(np.abs(rho - rho_o) / np.max(np.abs(rho_o))) < 1e-11


def test_generic_layout_matches():
    """A shifted SCF component takes the generic evaluator path; it must agree
    with the oracle as well."""
    from gala_amd.potential import SCFPotential
    S, T = workloads.scf_coefficients(nmax=4, lmax=3, seed=2)
    q0 = np.array([0.5, -0.2, 0.1])
    pot = SCFPotential(m=1e11, r_s=5.0, Snlm=S, Tnlm=T, origin=q0)
    orc = [O.scf(1e11, 5.0, S, T, q0=q0)]
    q = np.random.default_rng(1).uniform(-30, 30, (3, 500))
    g, go = pot.gradient(q), O.gradient(orc, q)
    assert np.max(np.abs(g - go) / np.linalg.norm(go, axis=0)) < 1e-11


def test_scf_leapfrog_and_dop853_parity():
    from gala_amd import _device as dev
    from gala_amd.integrate import dop853, leapfrog
    S, T = workloads.scf_coefficients()
    pot = product_scf(1e12, 20.0, S, T)
    orc = [O.scf(1e12, 20.0, S, T)]

    def vc(q):
        g = O.gradient(orc, q)
        r = np.linalg.norm(q, axis=0)
        return np.sqrt(r * np.abs(np.sum(g * q / r, axis=0)))
    w0, dt, _ = workloads.scf_leapfrog(vc, n=256, seed=3)
    t = np.arange(401) * dt
    ref = O.leapfrog(orc, w0, t, save_all=False)
    _, wf = leapfrog(pot, w0, dt, 400, save=False)
    assert np.max(state_rel_err(dev.to_host(wf), ref)) < 1e-10
    res = dop853(pot, w0[:, :64], 0.0, 300.0, 1.0, rtol=1e-8, atol=1e-10)
    ref2, st, _ = O.dop853_orbits(orc, w0[:, :64], 0.0, 300.0, 1.0, atol=1e-10, rtol=1e-8)
    assert np.array_equal(dev.to_host(res["status"]), st)
    assert np.median(state_rel_err(dev.to_host(res["w_final"]), ref2)) < 1e-9
