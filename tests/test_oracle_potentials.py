"""Pins the oracle's potentials (value, gradient, Hessian, density) with
closed forms derived by sympy from each potential's defining expression
(gala's to_sympy forms, builtin/core.py:237-242, 267-272, 556-563, 745-756),
the density from Poisson's equation (Laplacian / 4 pi G), and special
cases the reference code fixes."""
import numpy as np
import pytest
import sympy as sy

from oracle import oracle as O

G = O.G_GALACTIC
x, y, z = sy.symbols("x y z", real=True)
r = sy.sqrt(x ** 2 + y ** 2 + z ** 2)


def closed(kind):
    if kind == "hernquist":
        m, c = 5e9, 1.0
        return O.hernquist(m, c), -G * m / (r + c)
    if kind == "plummer":
        m, b = 1e5, 0.01
        return O.plummer(m, b), -G * m / sy.sqrt(r ** 2 + b ** 2)
    if kind == "miyamotonagai":
        m, a, b = 6.8e10, 3.0, 0.28
        return O.miyamotonagai(m, a, b), -G * m / sy.sqrt(x ** 2 + y ** 2 + (a + sy.sqrt(z ** 2 + b ** 2)) ** 2)
    if kind == "nfw":
        m, rs = 5.4e11, 15.6
        return O.nfw(m, rs), -G * m / rs * sy.log(1 + r / rs) / (r / rs)
    raise KeyError(kind)


KINDS = ["hernquist", "plummer", "miyamotonagai", "nfw"]


@pytest.mark.parametrize("kind", KINDS)
def test_closed_forms(kind):
    comp, expr = closed(kind)
    grad = [sy.diff(expr, v) for v in (x, y, z)]
    hess = [[sy.diff(g, v) for v in (x, y, z)] for g in grad]
    lap = sum(hess[i][i] for i in range(3))
    f_val = sy.lambdify((x, y, z), expr, "mpmath")
    f_grad = sy.lambdify((x, y, z), grad, "mpmath")
    f_hess = sy.lambdify((x, y, z), hess, "mpmath")
    f_lap = sy.lambdify((x, y, z), lap, "mpmath")
    import mpmath
    mpmath.mp.dps = 40
    rng = np.random.default_rng(KINDS.index(kind))
    scale = 0.05 if kind == "plummer" else 20.0
    q = rng.uniform(-scale, scale, (3, 40))
    val = O.value([comp], q)
    g = O.gradient([comp], q)
    h = O.hessian([comp], q)
    rho = O.density([comp], q)
    for i in range(q.shape[1]):
        args = [mpmath.mpf(float(v)) for v in q[:, i]]
        assert abs(val[i] - float(f_val(*args))) <= 1e-14 * abs(val[i])
        gr = np.array([float(v) for v in f_grad(*args)])
        assert np.max(np.abs(g[:, i] - gr)) <= 1e-13 * np.max(np.abs(gr))
        hr = np.array([[float(v) for v in row] for row in f_hess(*args)])
        assert np.max(np.abs(h[:, :, i] - hr)) <= 1e-12 * np.max(np.abs(hr))
        rho_ref = float(f_lap(*args)) / (4 * np.pi * G)
        assert abs(rho[i] - rho_ref) <= 1e-10 * abs(rho_ref)


def test_composite_is_sum_and_shift_is_translation():
    mw = O.milky_way_like()
    rng = np.random.default_rng(0)
    q = rng.uniform(-20, 20, (3, 50))
    total = sum(O.gradient([c], q) for c in mw)
    np.testing.assert_allclose(O.gradient(mw, q), total, rtol=1e-15, atol=1e-30)
    q0 = np.array([1.0, 2.0, -0.5])
    shifted = [O.hernquist(5e9, 1.0, q0=q0)]
    np.testing.assert_array_equal(O.gradient(shifted, q), O.gradient([O.hernquist(5e9, 1.0)],
                                                                     q - q0[:, None]))


def test_rotation_equivariance():
    """grad of a rotated component = R^T grad(R q) (potential_helpers.py:519-546)."""
    from tests.helpers import rotation
    R = rotation(3)
    base = O.miyamotonagai(6.8e10, 3.0, 0.28)
    rot = O.miyamotonagai(6.8e10, 3.0, 0.28, R=R)
    q = np.random.default_rng(1).uniform(-10, 10, (3, 30))
    np.testing.assert_allclose(O.gradient([rot], q), R.T @ O.gradient([base], R @ q), rtol=1e-12)
    np.testing.assert_allclose(O.value([rot], q), O.value([base], R @ q), rtol=1e-14)


def test_nfw_origin_special_cases():
    comp = [O.nfw(5.4e11, 15.6)]
    q = np.zeros((3, 1))
    assert O.value(comp, q)[0] == -G * 5.4e11 / 15.6
    assert np.all(np.isnan(O.gradient(comp, q)))


def test_d2_dr2_of_hernquist():
    """central second difference with h = 1e-2 approaches the analytic
    d2Phi/dr2 = -2 G m / (r + c)^3 to O(h^2)."""
    m, c = 5e9, 1.0
    q = np.array([[3.0], [4.0], [0.0]])  # r = 5
    d2 = O.d2_dr2([O.hernquist(m, c)], q)[0]
    exact = -2 * G * m / (5.0 + c) ** 3
    assert abs(d2 - exact) < 1e-4 * abs(exact)
