"""GPU leapfrog (gamd_leapfrog) against the oracle's restatement of
leapfrog_integrate_hamiltonian.

Bar: positions and velocities within 1e-10 relative after N steps (the
stated tolerance for Leapfrog).  Orbits whose own sensitivity to a one-ulp
change of the initial conditions exceeds that are gated at 100x their
measured sensitivity instead; the test reports how many there are."""
import numpy as np
import pytest

from gala_amd import workloads
from oracle import oracle as O
from tests.helpers import mw_oracle, mw_product, state_rel_err, vcirc_oracle

pytestmark = pytest.mark.gpu


def _sensitivity_tolerance(orc, w0, t):
    ref = O.leapfrog(orc, w0, t, save_all=False)
    pert = w0.copy()
    pert[0] = np.nextafter(pert[0], np.inf)
    alt = O.leapfrog(orc, pert, t, save_all=False)
    sens = state_rel_err(alt, ref)
    return ref, np.maximum(1e-10, 100.0 * sens)


def test_composite_subset_parity():
    from gala_amd.integrate import leapfrog
    from gala_amd import _device as dev
    orc = mw_oracle()
    w0, dt, _ = workloads.composite_leapfrog(vcirc_oracle(orc), n=1024, seed=0)
    nsteps = 2000
    t = np.arange(nsteps + 1) * dt
    ref, tol = _sensitivity_tolerance(orc, w0, t)
    wout, wfin = leapfrog(mw_product(), w0, dt, nsteps, save_stride=100)
    w = dev.to_host(wfin)
    err = state_rel_err(w, ref)
    loose = np.sum(tol > 1e-10)
    print(f"max rel err {err.max():.3e}; {loose} orbit(s) gated by their own sensitivity")
    assert np.all(err <= tol)
    assert loose <= 0.02 * len(err)
    # saved grid: step 0 is w0, and the saved step 2000 equals the final state bitwise
    wo = dev.to_host(wout)
    assert np.array_equal(wo[:, 0], w0)
    assert np.array_equal(wo[:, -1], w)


def test_save_all_trajectory_matches_oracle():
    from gala_amd.potential import Hamiltonian
    orc = mw_oracle()
    rng = np.random.default_rng(7)
    w0 = np.vstack([rng.uniform(5, 15, (3, 16)), rng.normal(0, 0.15, (3, 16))])
    t = np.arange(301) * 0.5
    ref = O.leapfrog(orc, w0, t, save_all=True)
    orbit = Hamiltonian(mw_product()).integrate_orbit(w0, t=t, Integrator="leapfrog")
    assert orbit.xyz.shape == (3, 301, 16)
    w = orbit.w()
    assert np.max(state_rel_err(w, ref)) < 1e-11
    assert np.array_equal(orbit.t, t)


def test_backward_and_single_orbit_squeeze():
    pot = mw_product()
    w0 = np.array([8.0, 0.0, 0.5, 0.0, 0.22, 0.01])
    fwd = pot.integrate_orbit(w0, dt=-1.0, n_steps=500)
    assert fwd.xyz.shape == (3, 501)
    assert fwd.t[-1] == -500.0
    ref = O.leapfrog(mw_oracle(), w0[:, None], fwd.t, save_all=True)[:, :, 0]
    assert np.max(state_rel_err(fwd.w(), ref)) < 1e-11
    last = pot.integrate_orbit(w0, dt=-1.0, n_steps=500, save_all=False)
    assert last.xyz.shape == (3, 1)
    assert np.array_equal(last.xyz[:, 0], fwd.xyz[:, -1])


def test_reversibility_and_energy_full_size():
    """Full 1024 x 10,000 workload: bounded energy error and time
    reversibility (size-independent properties of the symplectic map)."""
    from gala_amd.integrate import leapfrog
    from gala_amd import _device as dev
    from gala_amd.potential import Hamiltonian
    pot = mw_product()
    orc = mw_oracle()
    w0, dt, nsteps = workloads.composite_leapfrog(vcirc_oracle(orc), n=1024, seed=0)
    _, wfin = leapfrog(pot, w0, dt, nsteps, save=False)
    w1 = dev.to_host(wfin)
    H = Hamiltonian(pot)
    E0, E1 = H.energy(w0), H.energy(w1)
    dE = np.abs((E1 - E0) / E0)
    assert np.median(dE) < 1e-5
    back = w1.copy()
    back[3:] = -back[3:]
    _, wb = leapfrog(pot, back, dt, nsteps, save=False)
    wb = dev.to_host(wb)
    wb[3:] = -wb[3:]
    err = state_rel_err(wb, w0)
    assert np.median(err) < 1e-9


def test_uneven_lengths():
    from gala_amd.integrate import leapfrog
    from gala_amd import _device as dev
    orc = mw_oracle()
    rng = np.random.default_rng(11)
    n = 200
    w0 = np.vstack([rng.uniform(5, 15, (3, n)), rng.normal(0, 0.15, (3, n))])
    lens = rng.integers(0, 400, n)
    _, wfin = leapfrog(mw_product(), w0, 1.0, 0, save=False, n_steps_per_orbit=lens)
    w = dev.to_host(wfin)
    for L in (0, 1, 17, 399):
        sel = np.nonzero(lens == L)[0]
        if len(sel) == 0:
            continue
        ref = O.leapfrog(orc, w0[:, sel], np.arange(L + 1) * 1.0, save_all=False) if L > 0 \
            else w0[:, sel]
        assert np.max(state_rel_err(w[:, sel], ref)) < 1e-11


def test_ruth4_composite_parity():
    """gamd_ruth4 against the oracle's Ruth4 (ruth4.pyx:24-106) on the
    1024-orbit composite workload, 1e-10 relative or 100x the orbit's own
    one-ulp sensitivity; the saved grid's last step equals the final state."""
    from gala_amd.integrate import ruth4
    from gala_amd import _device as dev
    orc = mw_oracle()
    w0, dt, _ = workloads.composite_leapfrog(vcirc_oracle(orc), n=1024, seed=0)
    nsteps = 1000
    t = np.arange(nsteps + 1) * dt
    ref = O.ruth4(orc, w0, t, save_all=False)
    pert = w0.copy()
    pert[0] = np.nextafter(pert[0], np.inf)
    tol = np.maximum(1e-10, 100.0 * state_rel_err(O.ruth4(orc, pert, t, save_all=False), ref))
    wout, wfin = ruth4(mw_product(), w0, dt, nsteps, save_stride=50)
    w = dev.to_host(wfin)
    err = state_rel_err(w, ref)
    print(f"ruth4 max rel err {err.max():.3e}; {np.sum(tol > 1e-10)} orbit(s) gated")
    assert np.all(err <= tol)
    wo = dev.to_host(wout)
    assert np.array_equal(wo[:, 0], w0)
    assert np.array_equal(wo[:, -1], w)


def test_ruth4_integrate_orbit_save_all():
    from gala_amd.potential import Hamiltonian
    orc = mw_oracle()
    rng = np.random.default_rng(8)
    w0 = np.vstack([rng.uniform(5, 15, (3, 16)), rng.normal(0, 0.15, (3, 16))])
    t = np.arange(201) * 0.5
    ref = O.ruth4(orc, w0, t, save_all=True)
    orbit = Hamiltonian(mw_product()).integrate_orbit(w0, t=t, Integrator="ruth4")
    assert orbit.xyz.shape == (3, 201, 16)
    assert np.max(state_rel_err(orbit.w(), ref)) < 1e-11


@pytest.mark.parametrize("graph", [False, True])
def test_streamed_output_matches_device_output(graph):
    """gamd_leapfrog_stream (saved rows computed in chunks and copied to
    pinned host memory while the next chunk runs, optionally as one captured
    HIP graph) gives gamd_leapfrog's saved trajectory and final state bit for
    bit: chunk sizes from one row to all rows, a step count that is not a
    multiple of the stride, zero steps, the Milky-Way evaluator and the
    generic component loop.  The graph path is run three times in a row with
    other GPU work and synchronisations between, every run checked."""
    from gala_amd import _device as dev
    from gala_amd.integrate import leapfrog, leapfrog_to_host
    from gala_amd.potential import HernquistPotential, PlummerPotential
    orc = mw_oracle()
    w0, dt, _ = workloads.composite_leapfrog(vcirc_oracle(orc), n=700, seed=4)
    generic = HernquistPotential(m=1e11, c=2.0) + PlummerPotential(m=1e10, b=1.0,
                                                                   origin=[1.0, 0.5, -0.2])
    print(f"streamed leapfrog, graph={graph}", flush=True)
    for pot in (mw_product(), generic):
        for n_steps, stride in ((1000, 3), (0, 1), (5, 1)):
            wout, wfin = leapfrog(pot, w0, dt, n_steps, save_stride=stride)
            ref, ref_fin = dev.to_host(wout), dev.to_host(wfin)
            nsave = ref.shape[1]
            for rows in sorted({1, 7, 50, nsave}):
                reps = 3 if (graph and rows == 7) else 1
                for _ in range(reps):
                    got, gfin = leapfrog_to_host(pot, w0, dt, n_steps, stride, chunk_rows=rows,
                                                 graph=graph)
                    assert np.array_equal(got, ref), (n_steps, stride, rows)
                    assert np.array_equal(dev.to_host(gfin), ref_fin), (n_steps, stride, rows)
                    leapfrog(pot, w0, dt, 10, save=False)  # other work between the runs
                    dev.torch().cuda.synchronize()


def test_integrate_orbit_leapfrog_streams_to_host():
    """Hamiltonian.integrate_orbit's Leapfrog (gala's default integrator)
    returns the streamed trajectory: identical to the device path's."""
    from gala_amd import _device as dev
    from gala_amd.integrate import leapfrog
    from gala_amd.potential import Hamiltonian
    orc = mw_oracle()
    w0, dt, _ = workloads.composite_leapfrog(vcirc_oracle(orc), n=300, seed=6)
    t = np.arange(801) * dt
    orbit = Hamiltonian(mw_product()).integrate_orbit(w0, t=t, Integrator="leapfrog")
    wout, _ = leapfrog(mw_product(), w0, dt, 800)
    assert np.array_equal(orbit.w(), dev.to_host(wout))
