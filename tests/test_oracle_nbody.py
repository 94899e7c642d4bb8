"""The oracle's N-body restatements (direct_nbody_dop853,
leapfrog_integrate_nbody, ruth4_integrate_nbody, mock streams with
perturbers) pinned three ways:

* bit for bit against gala's own compiled dop853() with Fwrapper_direct_nbody
  and c_nbody_acceleration, and against gala's compiled c_gradient +
  c_nbody_gradient_symplectic inside the leapfrog loop (oracle/_ref, built
  from the reference tree where it is present);
* special cases that reduce to already-pinned paths: one massive body is a
  test-particle orbit; a massless perturber leaves a leapfrog stream unchanged;
* invariants of the two-body problem: total momentum (a linear invariant,
  kept by every Runge-Kutta method to rounding) and energy.
"""
import numpy as np
import pytest

from oracle import oracle as O

G = O.G_GALACTIC


def _bodies():
    """A progenitor-like Plummer sphere, an LMC-like composite body (Hernquist
    plus an offset, rotated Plummer component, exercising do_shift_rotate = 1
    with the component's own R kept and its q0 replaced by the body)."""
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    return [[O.plummer(2e8, 0.5)],
            [O.hernquist(1e11, 5.0), O.plummer(1e10, 2.0, q0=[1.0, 0.0, 0.0], R=R)]]


def _w0(n_test=2, seed=0):
    rng = np.random.default_rng(seed)
    w = [[15.0, 0.0, 1.0, 0.0, 0.17, 0.02],
         [-1.0, -40.0, -28.0, 0.06, -0.23, 0.21]]
    for _ in range(n_test):
        w.append(np.array(w[0]) + rng.normal(0, 1, 6) * [0.3, 0.3, 0.3, 0.005, 0.005, 0.005])
    return np.array(w, dtype=np.float64)


needs_ref = pytest.mark.skipif(O.ref_lib() is None, reason="reference build unavailable")


@needs_ref
@pytest.mark.parametrize("save_all", [True, False])
def test_nbody_dop853_bitwise_against_reference(save_all):
    ext = O.milky_way_like()
    w0 = _w0()
    t = np.linspace(0.0, 800.0, 201)
    a, ida, _ = O.nbody_dop853(ext, _bodies(), w0, t, atol=1e-10, rtol=1e-10,
                               save_all=save_all)
    b, idb, _ = O.nbody_dop853(ext, _bodies(), w0, t, atol=1e-10, rtol=1e-10,
                               save_all=save_all, ref=True)
    assert ida == idb == 1
    assert np.array_equal(a, b)
    # the massive bodies really interact: without them the test particles
    # follow other orbits
    c, _, _ = O.nbody_dop853(ext, [None, None], w0, t, save_all=save_all)
    assert not np.allclose(a[..., 2:, :], c[..., 2:, :], rtol=1e-6, atol=0)


@needs_ref
def test_nbody_leapfrog_bitwise_against_reference():
    ext = O.milky_way_like()
    w0 = _w0(3)
    t = np.arange(0.0, 601.0, 1.0)
    a = O.nbody_leapfrog(ext, _bodies(), w0, t)
    b = O.nbody_leapfrog(ext, _bodies(), w0, t, ref=True)
    assert np.array_equal(a, b)
    af = O.nbody_leapfrog(ext, _bodies(), w0, t, save_all=False)
    assert np.array_equal(af, a[-1])


@needs_ref
def test_nbody_ruth4_bitwise_against_reference():
    ext = O.milky_way_like()
    w0 = _w0(2)
    t = np.arange(0.0, 401.0, 1.0)
    a = O.nbody_ruth4(ext, _bodies(), w0, t)
    b = O.nbody_ruth4(ext, _bodies(), w0, t, ref=True)
    assert np.array_equal(a, b)
    assert np.array_equal(O.nbody_ruth4(ext, _bodies(), w0, t, save_all=False), a[-1])


@needs_ref
def test_mockstream_dop853_with_perturber_bitwise():
    """Release systems of progenitor + perturber + particles (nbody = 2)."""
    ext = O.milky_way_like()
    bodies = _bodies()
    rng = np.random.default_rng(1)
    nrel = 12
    t1 = -240.0 + 20.0 * np.arange(nrel)
    ns = np.full(nrel, 2, dtype=np.int32)
    ns[0] = 0
    ns[5] = 3
    bw = np.repeat(_w0(0)[None], nrel, axis=0)  # (nrel, 2, 6)
    bw[:, 0, 0] += 0.01 * np.arange(nrel)
    sw = np.repeat(bw[:, 0], ns, axis=0) + rng.normal(0, 0.05, (ns.sum(), 6)) * \
        np.array([1, 1, 1, 0.01, 0.01, 0.01])
    a = O.mockstream_dop853(ext, bodies, bw, t1, ns, sw, 0.0, 10.0)
    b = O.mockstream_dop853(ext, bodies, bw, t1, ns, sw, 0.0, 10.0, ref=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2]) and a[2][0] == 0 and np.all(a[2][1:] == 1)
    assert a[1].shape == (2, 6)


def test_one_massive_body_is_an_orbit():
    """One body alone: direct_nbody_dop853 is dop853 of one orbit (nbatch 1,
    stiffness test off) and leapfrog_integrate_nbody is the orbit leapfrog."""
    ext = O.milky_way_like()
    w0 = _w0(0)[:1]
    t = np.linspace(0.0, 500.0, 51)
    a, ida, _ = O.nbody_dop853(ext, [_bodies()[0]], w0, t)
    b, st = O.dop853(ext, w0.T, t, nstiff=-1, nbatch=1)[:2]
    assert ida == st[0] == 1
    assert np.array_equal(a[:, 0, :], b[:, :, 0].T)
    tl = np.arange(0.0, 301.0, 1.0)
    c = O.nbody_leapfrog(ext, [_bodies()[0]], w0, tl)
    d = O.leapfrog(ext, w0.T, tl)
    assert np.array_equal(c[:, 0, :], d[:, :, 0].T)
    e = O.nbody_ruth4(ext, [_bodies()[0]], w0, tl)
    f = O.ruth4(ext, w0.T, tl)
    assert np.array_equal(e[:, 0, :], f[:, :, 0].T)


def test_massless_perturber_leaves_leapfrog_stream_unchanged():
    """A perturber of zero mass adds exactly zero to every force, so the
    mock-stream leapfrog with it equals the one without, bit for bit; and
    the perturber itself moves as a test particle."""
    ext = O.milky_way_like()
    prog = [O.plummer(1e5, 0.01)]
    pert = [O.plummer(0.0, 1.0)]
    full_t = np.arange(0.0, 201.0, 1.0)
    bw0 = _w0(0)
    rng = np.random.default_rng(2)
    rel_idx = np.arange(0, 201, 10)
    ns = np.full(len(rel_idx), 2, dtype=np.int32)
    traj = O.leapfrog(ext, bw0[:1].T, full_t)[:, :, 0].T
    sw = np.repeat(traj[rel_idx], 2, axis=0) + rng.normal(0, 0.05, (2 * len(rel_idx), 6)) * \
        np.array([1, 1, 1, 0.01, 0.01, 0.01])
    pf1, s1 = O.mockstream_leapfrog(ext, prog, full_t, bw0[:1], full_t[rel_idx], ns, sw,
                                    full_t[-1])
    pf2, s2 = O.mockstream_leapfrog(ext, [prog, pert], full_t, bw0, full_t[rel_idx], ns, sw,
                                    full_t[-1])
    assert np.array_equal(s1, s2)
    assert np.array_equal(pf1, pf2[0])
    # with mass, the perturber changes the stream
    heavy = [O.plummer(1e11, 1.0)]
    _, s3 = O.mockstream_leapfrog(ext, [prog, heavy], full_t, bw0, full_t[rel_idx], ns, sw,
                                  full_t[-1])
    assert not np.allclose(s1, s3, rtol=1e-6, atol=0)


def test_two_body_invariants():
    """Two Plummer bodies, no external potential: DOP853 keeps the total
    momentum m1 v1 + m2 v2 to rounding and the energy to the tolerance."""
    m1, m2, b = 1e10, 3e9, 0.1
    bodies = [[O.plummer(m1, b)], [O.plummer(m2, b)]]
    v = np.sqrt(G * (m1 + m2) / 10.0)
    w0 = np.array([[-10 * m2 / (m1 + m2), 0, 0, 0, -0.7 * v * m2 / (m1 + m2), 0.01],
                   [10 * m1 / (m1 + m2), 0, 0, 0, 0.7 * v * m1 / (m1 + m2), -0.01 * m1 / m2]])
    t = np.linspace(0.0, 2000.0, 101)
    out, idid, _ = O.nbody_dop853([], bodies, w0, t, atol=1e-13, rtol=1e-13)
    assert idid == 1
    p = m1 * out[:, 0, 3:] + m2 * out[:, 1, 3:]
    pscale = m1 * np.abs(w0[0, 3:]).max()
    assert np.abs(p - p[0]).max() < 1e-12 * pscale

    def energy(w):
        r = np.linalg.norm(w[:, 0, :3] - w[:, 1, :3], axis=1)
        kin = 0.5 * m1 * (w[:, 0, 3:] ** 2).sum(1) + 0.5 * m2 * (w[:, 1, 3:] ** 2).sum(1)
        return kin - G * m1 * m2 / np.sqrt(r * r + b * b)
    E = energy(out)
    assert np.abs(E / E[0] - 1).max() < 1e-9
    # the orbit is bound and the bodies really move around each other
    r = np.linalg.norm(out[:, 0, :3] - out[:, 1, :3], axis=1)
    assert E[0] < 0 and r.min() < 0.9 * r.max()
