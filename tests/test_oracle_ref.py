"""The oracle against gala's own compiled C++ (oracle/_ref: cpotential.cpp,
chamiltonian.cpp, cframe.cpp, builtin_frames.cpp, dop853.cpp built from the
reference tree, see oracle/ref_driver.cpp).  Both are built with IEEE
arithmetic (-O2 -ffp-contract=off), so agreement is bit for bit."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import rotation

pytestmark = pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref not built")


def composites():
    R1, R2 = rotation(1), rotation(2)
    return {
        "mw": O.milky_way_like(),
        "shifted": [O.hernquist(3e10, 2.0), O.plummer(1e9, 0.5, q0=[1.0, -2.0, 0.5], R=R1),
                    O.miyamotonagai(5e10, 2.5, 0.3, R=R2), O.nfw(1e12, 20.0, q0=[0.1, 0.2, 0.3])],
        "plummer": [O.plummer(1e5, 0.01)],
    }


@pytest.mark.parametrize("name", ["mw", "shifted", "plummer"])
def test_evaluation_bitwise(name):
    comps = composites()[name]
    rng = np.random.default_rng(3)
    q = rng.uniform(-30, 30, (3, 300))
    for f in (O.value, O.gradient, O.density, O.d2_dr2):
        a, b = f(comps, q), f(comps, q, ref=True)
        assert np.array_equal(a, b), f.__name__
    if name != "shifted":  # gala does not rotate Hessians back (cpotential.cpp:309-311)
        assert np.array_equal(O.hessian(comps, q), O.hessian(comps, q, ref=True))


@pytest.mark.parametrize("nbatch,save_all", [(1, True), (1, False), (3, True), (100, False)])
def test_dop853_bitwise(nbatch, save_all):
    comps = O.milky_way_like()
    rng = np.random.default_rng(nbatch)
    n = 9
    w0 = np.vstack([rng.uniform(2, 25, (3, n)), rng.normal(0, 0.15, (3, n))])
    t = np.linspace(0.0, 800.0, 201)
    a, st_a, _ = O.dop853(comps, w0, t, atol=1e-10, rtol=1e-9, nbatch=nbatch, save_all=save_all)
    b, st_b = O.dop853(comps, w0, t, atol=1e-10, rtol=1e-9, nbatch=nbatch, save_all=save_all,
                       ref=True)
    assert np.array_equal(st_a, st_b)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("nbatch", [1, 4])
def test_dop853_hinit_bitwise(nbatch):
    """h0 = t[1] - t[0] = 0: dopcor starts from Hairer's initial step guess
    (hinit, dop853.cpp:18-86, called at :315-317), over the whole batch."""
    if O.ref_lib() is None:
        pytest.skip("gala's sources are not present")
    comps = O.milky_way_like()
    rng = np.random.default_rng(17)
    n = 8
    w0 = np.vstack([rng.uniform(2, 25, (3, n)), rng.normal(0, 0.15, (3, n))])
    for t in (np.array([0.0, 0.0, 700.0]), np.array([0.0, 0.0, -700.0])):
        a, st_a, stats = O.dop853(comps, w0, t, nbatch=nbatch, save_all=False)
        b, st_b = O.dop853(comps, w0, t, nbatch=nbatch, save_all=False, ref=True)
        assert np.all(st_a == 1) and np.array_equal(st_a, st_b)
        assert np.array_equal(a, b)


def test_dop853_backward_and_status_codes():
    comps = O.milky_way_like()
    w0 = np.array([[8.0], [0.1], [0.2], [0.01], [0.2], [0.02]])
    t = np.linspace(0.0, -500.0, 51)
    a, st_a, _ = O.dop853(comps, w0, t, nbatch=1)
    b, st_b = O.dop853(comps, w0, t, nbatch=1, ref=True)
    assert np.array_equal(a, b) and st_a[0] == st_b[0] == 1
    # nmax exceeded (dop853.cpp:340-347)
    a, st_a, _ = O.dop853(comps, w0, t, nbatch=1, nmax=3, save_all=False, atol=1e-13, rtol=1e-13)
    b, st_b = O.dop853(comps, w0, t, nbatch=1, nmax=3, save_all=False, atol=1e-13, rtol=1e-13,
                       ref=True)
    assert st_a[0] == st_b[0] == -2
    assert np.array_equal(a, b)


def test_mockstream_dop853_release_systems_bitwise():
    """The oracle's DOP853 release systems (progenitor + particles, one shared
    step, Fwrapper_direct_nbody forces, dop853_step's arguments) equal gala's
    own compiled dop853() driven as mockstream_dop853 does, bit for bit,
    including releases without particles and with one particle."""
    if O.ref_lib() is None:
        pytest.skip("reference build unavailable")
    ext = O.milky_way_like()
    prog = [O.plummer(1e5, 0.01)]
    rng = np.random.default_rng(0)
    nrel = 30
    t1 = -300.0 + 10.0 * np.arange(nrel)
    ns = np.full(nrel, 2, dtype=np.int32)
    ns[0] = 0
    ns[7] = 1
    ph = 0.01 * np.arange(nrel)
    pw = np.stack([15 * np.cos(ph), 15 * np.sin(ph), 0.3 + 0 * ph, -0.1 * np.sin(ph),
                   0.18 * np.cos(ph), 0.03 + 0 * ph], axis=1)
    sw = np.repeat(pw, ns, axis=0) + rng.normal(0, 0.05, (ns.sum(), 6)) * \
        np.array([1, 1, 1, 0.01, 0.01, 0.01])
    a = O.mockstream_dop853(ext, prog, pw, t1, ns, sw, 0.0, 10.0)
    b = O.mockstream_dop853(ext, prog, pw, t1, ns, sw, 0.0, 10.0, ref=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2]) and a[2][0] == 0 and np.all(a[2][1:] == 1)
