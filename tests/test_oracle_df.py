"""Pins of the oracle's FardalStreamDF._sample restatement (df.pyx:363-456
with get_rj_vj_R :61-92 and transform_from_sat :94-106) by what the
mathematics fixes, independent of the restatement's own formulas:

* on a circular Hernquist orbit the Jacobi radius is the closed form
  r_J = (G m / (Omega^2 - Phi''))^(1/3) with Omega^2 = G M / (r (r + c)^2)
  and Phi'' = -2 G M / (r + c)^3, and v_J = Omega r_J;
* the offsets are laid out in the satellite frame whose rows are
  (x_hat, L_hat x x_hat, L_hat): trailing particles at (+kx r_J, 0, +kz r_J)
  with velocity (0, +kvt kx v_J, +kvz v_J), leading ones mirrored, and the
  frame is orthonormal and turns with the orbit (a rotated orbit gives
  rotated offsets);
* the draws are consumed in the reference's order (per release time, the
  trailing block then the leading block, per particle kx, z, vt, vz), as a
  replay of successive RandomState.normal(mean, sd) calls gives them, with
  gala_modified's means and dispersions;
* release times with zero progenitor mass produce no particles and draw
  nothing.

The reference's own tests of the DF check only shapes and determinism
(tests/dynamics/mockstream/test_df.py); these checks are what pins it."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import rotation

G = O.G_GALACTIC
M, C = 1e11, 2.0          # host Hernquist
MSAT = 1e8                # progenitor mass


def circular(r, R=np.eye(3)):
    """Circular orbit in the host's Hernquist potential at radius r, in the
    x-y plane, turned by R: position, velocity, Omega, Phi''."""
    vc = np.sqrt(G * M * r) / (r + C)
    x = R @ np.array([r, 0.0, 0.0])
    v = R @ np.array([0.0, vc, 0.0])
    om2 = G * M / (r * (r + C) ** 2)
    d2 = -2.0 * G * M / (r + C) ** 3
    return x, v, np.sqrt(om2), d2


def sample(x, v, z, n=1, lead=True, trail=True, gala_modified=True, m=MSAT):
    ox, ov, ot = O.fardal_sample([O.hernquist(M, C)], G, x[None], v[None], np.array([0.0]),
                                 np.array([m]), np.array([n], dtype=np.int32), z, lead=lead,
                                 trail=trail, gala_modified=gala_modified)
    return ox - x, ov - v


@pytest.mark.parametrize("r", [3.0, 8.0, 25.0])
def test_jacobi_radius_closed_form_and_frame_axes(r):
    """kx = 1, all else zero: the trailing particle sits at r_J along x_hat
    with no velocity offset; the leading one at -r_J."""
    x, v, om, d2 = circular(r)
    rj = (G * MSAT / (om ** 2 - d2)) ** (1.0 / 3.0)
    # z for kx = 2 + 0.5 z0 = 1, z offset 0, vt = 0.3 + 0.5 z2 = 0, vz = 0
    one = [-2.0, 0.0, -0.6, 0.0]
    dx, dv = sample(x, v, np.array(one * 2))
    # c_d2_dr2 is a central difference with h = 1e-2 (cpotential.cpp:346-371):
    # Phi'' is exact to ~h^2 Phi''''/12, i.e. ~1e-5 relative here
    np.testing.assert_allclose(dx[0], [rj, 0.0, 0.0], rtol=2e-6, atol=1e-12 * rj)
    np.testing.assert_allclose(dx[1], [-rj, 0.0, 0.0], rtol=2e-6, atol=1e-12 * rj)
    assert np.max(np.abs(dv)) < 1e-12 * om * rj


def test_offsets_in_the_satellite_frame():
    """Each draw moves its own axis: kx along x_hat (and scales vt), kz along
    L_hat, kvt along L_hat x x_hat, kvz along L_hat; leading particles are
    mirrored."""
    r = 8.0
    x, v, om, d2 = circular(r)
    rj = (G * MSAT / (om ** 2 - d2)) ** (1.0 / 3.0)
    vj = om * rj
    z = np.array([0.4, -1.3, 0.7, 2.1])
    kx, kz, kvt, kvz = 2.0 + 0.5 * z[0], 0.5 * z[1], 0.3 + 0.5 * z[2], 0.5 * z[3]
    dx, dv = sample(x, v, np.concatenate([z, z]))
    for sgn, row in ((1.0, 0), (-1.0, 1)):
        np.testing.assert_allclose(dx[row], sgn * rj * np.array([kx, 0.0, kz]), rtol=2e-6,
                                   atol=1e-12 * rj)
        np.testing.assert_allclose(dv[row], sgn * vj * np.array([0.0, kvt * kx, kvz]),
                                   rtol=4e-6, atol=1e-12 * vj)
    # without gala_modified the tangential velocity is not scaled by kx and
    # the dispersions are Fardal's (0.4 for kx and kvt)
    dx2, dv2 = sample(x, v, z, lead=False, gala_modified=False)
    kx2, kvt2 = 2.0 + 0.4 * z[0], 0.3 + 0.4 * z[2]
    np.testing.assert_allclose(dx2[0], rj * np.array([kx2, 0.0, kz]), rtol=2e-6, atol=1e-12)
    np.testing.assert_allclose(dv2[0], vj * np.array([0.0, kvt2, kvz]), rtol=4e-6, atol=1e-12)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_frame_turns_with_the_orbit(seed):
    """The same circular orbit turned by a rotation Q: every offset is Q
    times the unrotated one (the frame's rows are x_hat, L_hat x x_hat,
    L_hat of the orbit itself, and it is orthonormal)."""
    Q = rotation(seed)
    z = np.random.default_rng(seed).standard_normal(4 * 6)
    x0, v0, _, _ = circular(6.0)
    x1, v1, _, _ = circular(6.0, Q)
    dx0, dv0 = sample(x0, v0, z, n=3)
    dx1, dv1 = sample(x1, v1, z, n=3)
    np.testing.assert_allclose(dx1, dx0 @ Q.T, rtol=0, atol=1e-10 * np.abs(dx0).max())
    np.testing.assert_allclose(dv1, dv0 @ Q.T, rtol=0, atol=1e-10 * np.abs(dv0).max())
    # lengths are preserved: the frame is orthonormal (to the rounding of
    # the central-difference Phi'', which moves r_J by ~1e-11 between the two)
    np.testing.assert_allclose(np.linalg.norm(dx1, axis=1), np.linalg.norm(dx0, axis=1),
                               rtol=1e-9)


@pytest.mark.parametrize("rs_kind", ["RandomState", "Generator"])
def test_draw_order_replays_random_state_normal(rs_kind):
    """Two release times with 2 and 3 particles: the offsets decode (in the
    satellite frame, by the closed-form r_J and v_J, so to the ~1e-6 of the
    central-difference Phi'') to the values of successive
    random_state.normal(mean, sd) calls made in the reference's
    loop order: time 0 trailing k = 0, 1, time 0 leading k = 0, 1, time 1
    trailing ..., each particle drawing kx, z, vt, vz."""
    seed = 42
    mk = (lambda: np.random.RandomState(seed)) if rs_kind == "RandomState" else \
        (lambda: np.random.default_rng(seed))
    counts = [2, 3]
    radii = [5.0, 12.0]
    xs, vs, rjs, vjs = [], [], [], []
    for r in radii:
        x, v, om, d2 = circular(r)
        rj = (G * MSAT / (om ** 2 - d2)) ** (1.0 / 3.0)
        xs.append(x)
        vs.append(v)
        rjs.append(rj)
        vjs.append(om * rj)
    ntot = 2 * sum(counts)
    z = np.asarray(mk().standard_normal(4 * ntot))
    ox, ov, ot = O.fardal_sample([O.hernquist(M, C)], G, np.array(xs), np.array(vs),
                                 np.array([10.0, 20.0]), np.array([MSAT, MSAT]),
                                 np.array(counts, dtype=np.int32), z)
    rs = mk()
    j = 0
    for i, n in enumerate(counts):
        for sgn in (1.0, -1.0):
            for k in range(n):
                kx = rs.normal(2.0, 0.5)
                kz = rs.normal(0.0, 0.5)
                kvt = rs.normal(0.3, 0.5)
                kvz = rs.normal(0.0, 0.5)
                dx = (ox[j] - xs[i]) / (sgn * rjs[i])
                dv = (ov[j] - vs[i]) / (sgn * vjs[i])
                np.testing.assert_allclose([dx[0], dx[2], dv[1] / kx, dv[2]],
                                           [kx, kz, kvt, kvz], rtol=5e-6, atol=1e-9)
                assert ot[j] == [10.0, 20.0][i]
                j += 1
    assert j == ntot


def test_zero_mass_release_is_skipped():
    """prog_m == 0 skips the release time (df.pyx:411-412): no particles and
    no draws, so the next release consumes the first draws."""
    x, v, _, _ = circular(8.0)
    z = np.array([0.3, -0.2, 0.5, 1.1] * 2)
    ox, ov, ot = O.fardal_sample([O.hernquist(M, C)], G, np.array([x, x]), np.array([v, v]),
                                 np.array([1.0, 2.0]), np.array([0.0, MSAT]),
                                 np.array([1, 1], dtype=np.int32), z)
    ref_x, ref_v = sample(x, v, z)
    np.testing.assert_array_equal(ox[:2] - x, ref_x)
    np.testing.assert_array_equal(ot[:2], [2.0, 2.0])
