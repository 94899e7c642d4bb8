"""Pins the oracle's fixed-step integrators (Leapfrog, leapfrog.pyx:24-121;
Ruth4, ruth4.pyx:24-106) by what the mathematics fixes, since gala stores
no trajectories for them and the Cython loops cannot be built here:

* convergence order on a circular orbit in a Hernquist sphere, whose exact
  solution is uniform rotation at Omega = v_c / r: halving dt divides the
  error by 4 for Leapfrog and by 16 for Ruth4 (a wrong Ruth coefficient or
  a dropped substep breaks fourth order);
* time reversibility: integrating forward, flipping the velocities and
  integrating again returns to the start (both schemes are palindromic);
* Ruth4's first substep has d_0 = 0, so starting from any state its first
  drift uses the initial velocity unchanged (checked through one step
  against a hand-written four-substep evaluation)."""
import numpy as np

from oracle import oracle as O

G = O.G_GALACTIC


def _circular(r0=8.0, m=1e11, c=2.0):
    GM = G * m
    vc = np.sqrt(GM * r0) / (r0 + c)
    return np.array([[r0], [0.0], [0.0], [0.0], [vc], [0.0]]), vc / r0


def _final_error(method, nsteps, periods=2.0):
    orc = [O.hernquist(1e11, 2.0)]
    w0, om = _circular()
    T = periods * 2 * np.pi / om
    t = np.linspace(0.0, T, nsteps + 1)
    w = method(orc, w0, t, save_all=False)
    exact = np.array([8.0 * np.cos(om * T), 8.0 * np.sin(om * T)])
    return np.hypot(w[0, 0] - exact[0], w[1, 0] - exact[1])


def test_convergence_orders():
    lf = [_final_error(O.leapfrog, n) for n in (400, 800)]
    r4 = [_final_error(O.ruth4, n) for n in (100, 200)]
    assert 3.6 < lf[0] / lf[1] < 4.4, lf
    assert 14.0 < r4[0] / r4[1] < 18.0, r4
    # fourth order with three evaluations per step beats second order with one
    assert r4[1] < lf[1]


def test_time_reversibility():
    orc = O.milky_way_like()
    rng = np.random.default_rng(4)
    w0 = np.vstack([rng.uniform(4, 12, (3, 8)), rng.normal(0, 0.2, (3, 8))])
    t = np.arange(201) * 0.5
    for method in (O.leapfrog, O.ruth4):
        w1 = method(orc, w0, t, save_all=False)
        back = w1.copy()
        back[3:] *= -1.0
        w2 = method(orc, back, t, save_all=False)
        w2[3:] *= -1.0
        assert np.max(np.abs(w2 - w0)) < 1e-10


def test_ruth4_one_step_by_hand():
    orc = O.milky_way_like()
    w0 = np.array([[7.0], [1.0], [0.4], [0.05], [0.2], [-0.02]])
    dt = 0.7
    a = 2.0 ** (1.0 / 3.0)
    cs = [1 / (2 * (2 - a)), (1 - a) / (2 * (2 - a)), (1 - a) / (2 * (2 - a)), 1 / (2 * (2 - a))]
    ds = [0.0, 1 / (2 - a), -a / (2 - a), 1 / (2 - a)]
    x, v = w0[:3, 0].copy(), w0[3:, 0].copy()
    for j in range(4):
        g = O.gradient(orc, x.reshape(3, 1))[:, 0]
        v = v - ds[j] * g * dt
        x = x + cs[j] * v * dt
    w = O.ruth4(orc, w0, np.array([0.0, dt]), save_all=False)
    np.testing.assert_allclose(w[:3, 0], x, rtol=1e-14)
    np.testing.assert_allclose(w[3:, 0], v, rtol=1e-14)
