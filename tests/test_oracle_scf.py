"""Pins the oracle's SCF restatement (gala potential/scf/src/bfe.cpp with
GSL replaced by recurrences) against:
  * the Fortran SCF golden vectors gala's tests hold (5 coefficient sets,
    rtol 1e-6 as tests/potential/scf/test_accp_fortran.py:77-156);
  * an independent scipy evaluation of the same series (scipy.special
    eval_gegenbauer / lpmv / gamma, as gala's test_bfe.py:65-184 pure_py);
  * the closed form it must reduce to: S_000 = 1 is a Hernquist sphere of
    mass M and scale r_s (test_bfe.py:34-62, test_class.py:27-57).
"""
import os
from math import factorial

import numpy as np
import pytest
from scipy.special import eval_gegenbauer, lpmv

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
NAMES = ["simple-hernquist", "multi-hernquist", "simple-nonsph", "random", "wang-zhao"]


def fortran_case(name):
    d = np.load(os.path.join(HERE, "golden", "scf_fortran.npz"))
    coeff = d[f"{name}/coeff"]
    nmax = int(coeff[:, 0].max())
    lmax = int(coeff[:, 1].max())
    S = np.zeros((nmax + 1, lmax + 1, lmax + 1))
    T = np.zeros_like(S)
    for n, l, m, cc, sc in coeff:
        n, l, m = int(n), int(l), int(m)
        if l != 0:  # H&O 1992 -> Lowing 2011 normalisation (test_accp_fortran.py:50-56)
            fac = np.sqrt(4 * np.pi) * np.sqrt((2 * l + 1) / (4 * np.pi) * factorial(l - m)
                                               / factorial(l + m))
            cc /= fac
            sc /= fac
        S[n, l, m] = cc
        T[n, l, m] = sc
    return d["xyz"].T.copy(), S, T, d[f"{name}/accp"]


@pytest.mark.parametrize("name", NAMES)
def test_fortran_golden(name):
    xyz, S, T, accp = fortran_case(name)
    comp = [O.scf(1.0, 1.0, S, T, G=1.0)]
    pot = O.value(comp, xyz)
    np.testing.assert_allclose(pot, -accp[:, -1], rtol=1e-6)
    grad = O.gradient(comp, xyz)
    np.testing.assert_allclose(grad.T, accp[:, :3], rtol=1e-6, atol=1e-7 * np.abs(accp[:, :3]).max())


def scipy_series(xyz, S, T):
    """Potential and gradient of the basis expansion (G = M = r_s = 1) with
    scipy special functions, term by term."""
    r = np.sqrt(np.sum(xyz ** 2, axis=0))
    X = xyz[2] / r
    sinth = np.sqrt(1 - X ** 2)
    phi = np.arctan2(xyz[1], xyz[0])
    xi = (r - 1) / (r + 1)
    nmax, lmax = S.shape[0] - 1, S.shape[1] - 1
    pot = 0.0
    sg = np.zeros_like(xyz)
    s4p = np.sqrt(4 * np.pi)
    for l in range(lmax + 1):
        for m in range(l + 1):
            norm = np.sqrt((2 * l + 1) / (4 * np.pi) * factorial(l - m) / factorial(l + m))
            Y = norm * lpmv(m, l, X)
            dY = norm / sinth * (l * X * lpmv(m, l, X) - (l + m) * lpmv(m, l - 1, X)) if l else 0.0
            for n in range(nmax + 1):
                a = S[n, l, m] * np.cos(m * phi) + T[n, l, m] * np.sin(m * phi)
                b = -S[n, l, m] * np.sin(m * phi) + T[n, l, m] * np.cos(m * phi)
                Cn = eval_gegenbauer(n, 2 * l + 1.5, xi)
                phinl = -s4p * r ** l * (1 + r) ** (-2 * l - 1) * Cn
                Cn1 = eval_gegenbauer(n - 1, 2 * l + 2.5, xi) if n > 0 else 0.0
                dphinl = s4p * r ** (l - 1) * (1 + r) ** (-3 - 2 * l) * (
                    -2 * (3 + 4 * l) * r * Cn1 + (1 + r) * (l * (r - 1) + r) * Cn)
                pot = pot + phinl * Y * a
                sg[0] += dphinl * Y * a
                sg[1] += phinl * dY * a / r
                sg[2] += m * phinl * Y * b / (r * sinth)
    cphi, sphi = np.cos(phi), np.sin(phi)
    g = np.array([sinth * cphi * sg[0] + X * cphi * sg[1] - sphi * sg[2],
                  sinth * sphi * sg[0] + X * sphi * sg[1] + cphi * sg[2],
                  X * sg[0] - sinth * sg[1]])
    return pot, g


def test_scipy_series_random_coefficients():
    rng = np.random.default_rng(4)
    nmax, lmax = 6, 4
    S = rng.uniform(-1, 1, (nmax + 1, lmax + 1, lmax + 1))
    T = rng.uniform(-1, 1, (nmax + 1, lmax + 1, lmax + 1))
    T[:, :, 0] = 0.0
    xyz = rng.uniform(-2, 2, (3, 128))
    pot, g = scipy_series(xyz, S, T)
    comp = [O.scf(1.0, 1.0, S, T, G=1.0)]
    np.testing.assert_allclose(O.value(comp, xyz), pot, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(O.gradient(comp, xyz), g, rtol=1e-10, atol=1e-12)


def test_hernquist_reduction():
    M, rs = 1e10, 3.5
    S = np.zeros((7, 3, 3))
    T = np.zeros_like(S)
    S[0, 0, 0] = 1.0
    rr = np.linspace(0.1, 10.0, 128)
    xyz = np.array([rr * np.cos(np.pi / 4) * np.sin(np.pi / 4),
                    rr * np.sin(np.pi / 4) * np.sin(np.pi / 4), rr * np.cos(np.pi / 4)])
    scf = [O.scf(M, rs, S, T)]
    hq = [O.hernquist(M, rs)]
    np.testing.assert_allclose(O.value(scf, xyz), O.value(hq, xyz), rtol=1e-13)
    np.testing.assert_allclose(O.gradient(scf, xyz), O.gradient(hq, xyz), rtol=1e-12)
    np.testing.assert_allclose(O.density(scf, xyz), O.density(hq, xyz), rtol=1e-12)


def test_gradient_matches_finite_difference_of_value():
    rng = np.random.default_rng(8)
    S = rng.normal(size=(5, 4, 4)) * 0.3
    T = rng.normal(size=(5, 4, 4)) * 0.3
    S[0, 0, 0] = 1.0
    comp = [O.scf(1.0, 1.0, S, T, G=1.0)]
    q = rng.uniform(-3, 3, (3, 20))
    g = O.gradient(comp, q)
    h = 1e-5
    fd = np.zeros_like(q)
    for k in range(3):
        e = np.zeros((3, 1))
        e[k] = h
        fd[k] = (O.value(comp, q + e) - O.value(comp, q - e)) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-8)
