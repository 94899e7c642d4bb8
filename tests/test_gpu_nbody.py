"""N-body systems on the GPU against the oracle (itself bit-for-bit equal to
gala's compiled dop853 + Fwrapper_direct_nbody and c_nbody_gradient_symplectic,
tests/test_oracle_nbody.py): DirectNBody with several massive bodies under
DOP853, Leapfrog and Ruth4, and mock streams with a massive perturber.

Tolerances: DOP853 states within 1e-8 relative, Leapfrog within 1e-10, or
100x the oracle's own sensitivity to a one-ulp change of the initial state
(close encounters with a massive body amplify ulp differences of the
force evaluation).
"""
import numpy as np
import pytest

from gala_amd import workloads
from oracle import oracle as O
from tests.helpers import mw_oracle, mw_product, state_rel_err

pytestmark = pytest.mark.gpu

TH = 0.3
ROT = np.array([[np.cos(TH), -np.sin(TH), 0], [np.sin(TH), np.cos(TH), 0], [0, 0, 1.0]])


def _lmc():
    """An LMC-like composite body: a Hernquist sphere plus an offset, rotated
    Plummer component (the body's position replaces the offset; R is kept)."""
    from gala_amd.potential import CompositePotential, HernquistPotential, PlummerPotential
    p = CompositePotential()
    p["h"] = HernquistPotential(m=1e11, c=5.0)
    p["p"] = PlummerPotential(m=1e10, b=2.0, origin=[1.0, 0.0, 0.0], R=ROT)
    return p, [O.hernquist(1e11, 5.0), O.plummer(1e10, 2.0, q0=[1.0, 0.0, 0.0], R=ROT)]


def _system(n_test, seed=0):
    """(w0 (n, 6): progenitor-like body, LMC-like body, n_test particles)."""
    rng = np.random.default_rng(seed)
    w = [[15.0, 0.0, 1.0, 0.0, 0.17, 0.02], [-1.0, -40.0, -28.0, 0.06, -0.23, 0.21]]
    for _ in range(n_test):
        w.append(np.array(w[0]) + rng.normal(0, 1, 6) * [0.5, 0.5, 0.5, 0.01, 0.01, 0.01])
    return np.array(w, dtype=np.float64)


def _gate(ref, alt, floor):
    """per-body tolerance: floor, or 100x the one-ulp sensitivity"""
    return np.maximum(floor, 100 * state_rel_err(alt, ref))


@pytest.mark.parametrize("save_all", [True, False])
def test_direct_nbody_dop853_rows_match_waves(save_all, monkeypatch):
    """Systems of at most 4 bodies run on one 16-lane row (a body per quad,
    an axis per lane) or on one wave (a body per lane): dense output, final
    states and status must agree bit for bit.  GAMD_DOP853_BODY_ROWS=0 forces
    the wave, =1 the row."""
    from gala_amd.dynamics import DirectNBody
    from gala_amd.potential import PlummerPotential
    lmc, _ = _lmc()
    w = _system(1, seed=3)
    pots = [PlummerPotential(m=2e8, b=0.5), lmc, None]
    t = np.linspace(0.0, 600.0, 121)
    res = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("GAMD_DOP853_BODY_ROWS", mode)
        nb = DirectNBody(w.T, pots, external_potential=mw_product(), save_all=save_all)
        res[mode] = nb.integrate_orbit(t=t).w()
    print(f"nbody dop853 rows vs wave: {res['0'].shape}")
    assert np.all(np.isfinite(res["0"]))
    assert np.array_equal(res["0"], res["1"])


@pytest.mark.parametrize("save_all", [True, False])
def test_direct_nbody_dop853_matches_oracle(save_all):
    from gala_amd.dynamics import DirectNBody
    from gala_amd.potential import PlummerPotential
    lmc, lmc_o = _lmc()
    w = _system(3)
    # a test particle first: DirectNBody puts the massive bodies in front
    order = [2, 0, 1, 3, 4]
    pots = [None, PlummerPotential(m=2e8, b=0.5), lmc, None, None]
    nb = DirectNBody(w[order].T, pots, external_potential=mw_product(), save_all=save_all)
    t = np.linspace(0.0, 800.0, 161)
    got = nb.integrate_orbit(t=t).w()
    reorg = w[[0, 1, 2, 3, 4]]
    bodies = [[O.plummer(2e8, 0.5)], lmc_o]
    ref, idid, _ = O.nbody_dop853(mw_oracle(), bodies, reorg, t, save_all=save_all)
    pert = reorg.copy()
    pert[:, 0] = np.nextafter(pert[:, 0], np.inf)
    alt = O.nbody_dop853(mw_oracle(), bodies, pert, t, save_all=save_all)[0]
    assert idid == 1
    ref_w, alt_w = np.moveaxis(ref, -1, 0), np.moveaxis(alt, -1, 0)  # (6, [ntimes,] n)
    err = state_rel_err(got, ref_w)
    tol = _gate(ref_w, alt_w, 1e-8)
    print(f"nbody dop853: max rel err {err.max():.3e}")
    assert got.shape == ref_w.shape
    assert np.all(err <= tol)


@pytest.mark.parametrize("n_massive,n_test", [(3, 64), (6, 5), (0, 3)])
def test_direct_nbody_leapfrog_matches_oracle(n_massive, n_test):
    """Leapfrog N-body (bodies move in index order within a step) with 3
    bodies per lane in registers, 6 (the spilled variant) and none."""
    from gala_amd.dynamics import DirectNBody
    from gala_amd.integrate import LeapfrogIntegrator
    from gala_amd.potential import HernquistPotential, PlummerPotential
    rng = np.random.default_rng(n_massive)
    w = _system(n_test + max(n_massive - 2, 0))
    w[2:n_massive, :3] += rng.normal(0, 5.0, (max(n_massive - 2, 0), 3))
    lmc, lmc_o = _lmc()
    pots, bodies = [], []
    for j in range(n_massive):
        if j == 1:
            pots.append(lmc)
            bodies.append(lmc_o)
        elif j % 2:
            pots.append(HernquistPotential(m=3e9, c=1.0))
            bodies.append([O.hernquist(3e9, 1.0)])
        else:
            pots.append(PlummerPotential(m=2e8, b=0.5))
            bodies.append([O.plummer(2e8, 0.5)])
    pots += [None] * (w.shape[0] - n_massive)
    nb = DirectNBody(w.T, pots, external_potential=mw_product())
    t = np.arange(0.0, 501.0, 1.0)
    got = nb.integrate_orbit(t=t, Integrator=LeapfrogIntegrator).w()
    ref = O.nbody_leapfrog(mw_oracle(), bodies, w, t)
    pert = w.copy()
    pert[:, 0] = np.nextafter(pert[:, 0], np.inf)
    alt = O.nbody_leapfrog(mw_oracle(), bodies, pert, t)
    ref_w, alt_w = np.moveaxis(ref, -1, 0), np.moveaxis(alt, -1, 0)
    err = state_rel_err(got, ref_w)
    tol = _gate(ref_w, alt_w, 1e-10)
    print(f"nbody leapfrog K={n_massive}: max rel err {err.max():.3e}")
    assert np.all(err <= tol)
    fin = DirectNBody(w.T, pots, external_potential=mw_product(), save_all=False)
    wf = fin.integrate_orbit(t=t, Integrator=LeapfrogIntegrator).w()
    assert np.array_equal(wf, got[:, -1])


@pytest.mark.parametrize("n_massive", [2, 5])
def test_direct_nbody_ruth4_matches_oracle(n_massive):
    """Ruth4 N-body (all four substeps of a body before the next body moves)
    with bodies in registers (2) and spilled (5)."""
    from gala_amd.dynamics import DirectNBody
    from gala_amd.integrate import Ruth4Integrator
    from gala_amd.potential import PlummerPotential
    rng = np.random.default_rng(7)
    w = _system(8 + n_massive - 2)
    w[2:n_massive, :3] += rng.normal(0, 5.0, (n_massive - 2, 3))
    lmc, lmc_o = _lmc()
    pots = [PlummerPotential(m=2e8, b=0.5), lmc] + [PlummerPotential(m=1e9, b=0.3)] * \
        (n_massive - 2)
    bodies = [[O.plummer(2e8, 0.5)], lmc_o] + [[O.plummer(1e9, 0.3)]] * (n_massive - 2)
    pots += [None] * (w.shape[0] - n_massive)
    t = np.arange(0.0, 401.0, 1.0)
    got = DirectNBody(w.T, pots, external_potential=mw_product()).integrate_orbit(
        t=t, Integrator=Ruth4Integrator).w()
    ref = O.nbody_ruth4(mw_oracle(), bodies, w, t)
    pert = w.copy()
    pert[:, 0] = np.nextafter(pert[:, 0], np.inf)
    alt = O.nbody_ruth4(mw_oracle(), bodies, pert, t)
    ref_w, alt_w = np.moveaxis(ref, -1, 0), np.moveaxis(alt, -1, 0)
    err = state_rel_err(got, ref_w)
    print(f"nbody ruth4 K={n_massive}: max rel err {err.max():.3e}")
    assert np.all(err <= _gate(ref_w, alt_w, 1e-10))


def _release_inputs(nrel=40, seed=4):
    """Bodies (progenitor, LMC-like) at nrel release times from the oracle's
    N-body DOP853, and 1..3 particles per release: systems of 3 and 4 bodies
    run on quads, those of 5 on waves."""
    lmc, lmc_o = _lmc()
    prog_w0, prog_mass, b, _, _ = workloads.mockstream_progenitor()
    w = np.array([prog_w0, [-1.0, -40.0, -28.0, 0.06, -0.23, 0.21]])
    time = -200.0 + 5.0 * np.arange(nrel)
    bodies = [[O.plummer(prog_mass, b)], lmc_o]
    bw, idid, _ = O.nbody_dop853(mw_oracle(), bodies, w, time)
    assert idid == 1
    rng = np.random.default_rng(seed)
    ns = rng.integers(1, 4, nrel).astype(np.int32)
    ns[0] = 0
    sw = np.repeat(bw[:, 0], ns, axis=0) + rng.normal(0, 1.0, (ns.sum(), 6)) * \
        np.array([0.05, 0.05, 0.05, 0.002, 0.002, 0.002])
    return lmc, bodies, prog_mass, b, time, np.ascontiguousarray(bw), ns, np.ascontiguousarray(sw)


def test_dop853_release_kernel_with_perturber():
    from gala_amd import _device as dev
    from gala_amd import _lib
    from gala_amd.potential import PlummerPotential
    lmc, bodies, prog_mass, b, time, bw, ns, sw = _release_inputs()
    ext = mw_product()
    prog = PlummerPotential(m=prog_mass, b=b)
    t = dev.torch()
    nrel, ntot, K = len(time), sw.shape[0], 2
    off = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
    out = t.empty((ntot, 6), dtype=t.float64, device="cuda")
    bf = t.empty((nrel, K, 6), dtype=t.float64, device="cuda")
    status = t.empty(nrel, dtype=t.int32, device="cuda")
    keep = [dev.to_dev(bw), dev.to_dev(time), dev.to_dev(ns, dtype=t.int32),
            dev.to_dev(off, dtype=t.int64), dev.to_dev(sw)]
    lib = _lib.load()
    print(f"releases {nrel}, particles {ntot}, systems of {K + ns.min()}..{K + ns.max()} bodies",
          flush=True)
    _lib.check(lib.gamd_mockstream_dop853(ext.c_handle, _lib.body_array([prog.c_handle,
                                                                         lmc.c_handle]), K,
                                          keep[0].data_ptr(), nrel, keep[1].data_ptr(),
                                          keep[2].data_ptr(), keep[3].data_ptr(),
                                          keep[4].data_ptr(), 0.0, 5.0, 1e-10, 1e-10, 0,
                                          out.data_ptr(), bf.data_ptr(), status.data_ptr(), None,
                                          dev.stream_ptr()), "gamd_mockstream_dop853")
    ref_out, ref_bf, ref_st, _ = O.mockstream_dop853(mw_oracle(), bodies, bw, time, ns, sw,
                                                     0.0, 5.0)
    pert = sw.copy()
    pert[:, 0] = np.nextafter(pert[:, 0], np.inf)
    alt = O.mockstream_dop853(mw_oracle(), bodies, bw, time, ns, pert, 0.0, 5.0)[0]
    st = dev.to_host(status)
    assert np.array_equal(st, ref_st) and st[0] == 0 and np.all(st[1:] == 1)
    err = state_rel_err(dev.to_host(out).T, ref_out.T)
    tol = _gate(ref_out.T, alt.T, 1e-8)
    print(f"release systems with perturber: max rel err {err.max():.3e}")
    assert np.all(err <= tol)
    last = int(np.nonzero(ns)[0][-1])
    assert np.all(state_rel_err(dev.to_host(bf)[last].T, ref_bf.T) < 1e-8)


@pytest.mark.parametrize("integrator", ["leapfrog", "dop853"])
def test_generator_with_perturber_against_oracle_pipeline(integrator):
    """MockStreamGenerator.run(nbody=DirectNBody(LMC-like body)) against the
    oracle pipeline fed the product's own DF samples: the bodies' orbits
    backwards, reversed, then the release systems (mockstream_generator.py:
    229-330; mockstream.pyx:176-303 for DOP853, 442-620 for Leapfrog)."""
    from gala_amd.dynamics import DirectNBody, FardalStreamDF, MockStreamGenerator, Orbit
    from gala_amd.potential import PlummerPotential
    lmc, lmc_o = _lmc()
    prog_w0, prog_mass, b, dt, _ = workloads.mockstream_progenitor()
    lmc_w0 = np.array([-1.0, -40.0, -28.0, 0.06, -0.23, 0.21])
    ext = mw_product()
    nsteps = 300

    def gen(seed):
        df = FardalStreamDF(gala_modified=True, random_state=np.random.RandomState(seed))
        return MockStreamGenerator(df, ext, progenitor_potential=PlummerPotential(m=prog_mass,
                                                                                  b=b))
    nbody = DirectNBody(lmc_w0[:, None], [lmc], external_potential=ext)
    stream, bodies_f = gen(21).run(prog_w0, prog_mass, nbody=nbody, Integrator=integrator,
                                   dt=dt, n_steps=nsteps)
    assert bodies_f.xyz.shape == (3, 2)
    orc = mw_oracle()
    bodies = [[O.plummer(prog_mass, b)], lmc_o]
    w0 = np.array([prog_w0, lmc_w0])
    t = np.arange(nsteps + 1) * dt
    # the bodies' backward orbits as run() computes them (DirectNBody with the
    # progenitor and the perturber, mockstream_generator.py:229-232), so the
    # DF samples below are the product's own; checked against the oracle
    sys = DirectNBody(w0.T, [PlummerPotential(m=prog_mass, b=b), lmc], external_potential=ext)
    back = np.moveaxis(sys.integrate_orbit(Integrator=integrator, t=t).w(), 0, -1)
    if integrator == "leapfrog":
        ref_back = O.nbody_leapfrog(orc, bodies, w0, t)
    else:
        ref_back, idid, _ = O.nbody_dop853(orc, bodies, w0, t)
        assert idid == 1
    assert np.all(state_rel_err(np.moveaxis(back, -1, 0), np.moveaxis(ref_back, -1, 0)) <
                  (1e-9 if integrator == "leapfrog" else 1e-8))
    back = back[::-1]  # (ntimes, K, 6), forward in time
    orbit_t = t[::-1]
    w = back[:, 0].T
    g = gen(21)
    ics = g.df.sample(Orbit(pos=w[:3], vel=w[3:], t=orbit_t, hamiltonian=g.hamiltonian),
                      prog_mass, hamiltonian=g.hamiltonian)
    rel = np.unique(ics.release_time)
    ns = np.full(len(rel), 2, dtype=np.int32)
    sw = np.ascontiguousarray(ics.w().T)
    pert = sw.copy()
    pert[:, 0] = np.nextafter(pert[:, 0], np.inf)
    if integrator == "leapfrog":
        run = lambda s, bw=back[0]: O.mockstream_leapfrog(orc, bodies, orbit_t, bw, rel, ns, s,
                                                          orbit_t[-1])
        ref_bf, ref_out = run(sw)
        # sensitivity to one ulp of the particle and of the bodies' states
        # (the GPU and CPU body trajectories differ at that level)
        bw2 = back[0].copy()
        bw2[:, 0] = np.nextafter(bw2[:, 0], np.inf)
        a1, a2 = run(pert)[1], run(sw, bw2)[1]
        alt = np.where(state_rel_err(a1.T, ref_out.T)[:, None] >
                       state_rel_err(a2.T, ref_out.T)[:, None], a1, a2)
        floor = 1e-10
    else:
        fwd, idid, _ = O.nbody_dop853(orc, bodies, back[0], rel)
        assert idid == 1
        run = lambda s: O.mockstream_dop853(orc, bodies, fwd, rel, ns, s, orbit_t[-1],
                                            rel[1] - rel[0])
        ref_out, ref_bf, ref_st, _ = run(sw)
        assert np.all(ref_st == 1)
        alt = run(pert)[0]
        floor = 1e-8
    err = state_rel_err(stream.w(), ref_out.T)
    tol = _gate(ref_out.T, alt.T, floor)
    print(f"{integrator} stream with perturber: median {np.median(err):.3e}, max {err.max():.3e}; "
          f"{np.sum(tol > floor)} of {len(tol)} particle(s) gated by their own sensitivity")
    assert np.all(err <= tol)
    assert np.mean(tol > floor) <= 0.10
    assert np.all(state_rel_err(bodies_f.w(), ref_bf.T) < floor * 100)
