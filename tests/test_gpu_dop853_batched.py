"""GPU DOP853 with gala's default batching (gamd_dop853_batched) against the
oracle's restatement of dop853_integrate_hamiltonian with nbatch orbits per
system, which tests/test_oracle_ref.py matches bit for bit against gala's own
compiled dop853.cpp.

Hamiltonian.integrate_orbit(..., Integrator='dopri853') passes nbatch=100
(chamiltonian.pyx:348): each block of 100 orbits is one 600-dimensional
system with one shared adaptive step (dop853.pyx:228-245), whose results
differ from per-orbit integration by up to ~1e-5.  Bars: identical status
codes, identical step counts per batch, dense output and final states within
1e-8 relative (the stated DOP853 tolerance) or, for an orbit whose own
one-ulp sensitivity in the oracle exceeds that, within 100x that
sensitivity; the number of such gated orbits is printed and bounded."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import mw_oracle, mw_product, state_rel_err

pytestmark = pytest.mark.gpu


def _ics(n, seed):
    rng = np.random.default_rng(seed)
    return np.vstack([rng.uniform(3, 20, (3, n)), rng.normal(0, 0.12, (3, n))])


def _sensitivity(orc, w0, t, nbatch, save_all, **kw):
    """Largest response of the oracle's result to a one-ulp change of x, y or
    v_z of every orbit (per orbit, over all saved times)."""
    ref = O.dop853(orc, w0, t, nbatch=nbatch, save_all=save_all, **kw)[0]
    sens = np.zeros(w0.shape[1])
    for k, direction in ((0, np.inf), (1, -np.inf), (5, np.inf)):
        pert = w0.copy()
        pert[k] = np.nextafter(pert[k], direction)
        alt = O.dop853(orc, pert, t, nbatch=nbatch, save_all=save_all, **kw)[0]
        e = state_rel_err(alt, ref)
        sens = np.maximum(sens, e.max(axis=0) if e.ndim == 2 else e)
    return sens


def _check(w, ref, sens, what, max_gated):
    err = state_rel_err(w, ref)
    err = err.max(axis=0) if err.ndim == 2 else err
    tol = np.maximum(1e-8, 100.0 * sens)
    gated = int(np.sum(tol > 1e-8))
    bad = np.nonzero(err > tol)[0]
    print(f"{what}: max err {err.max():.3e}, median {np.median(err):.3e}; "
          f"{gated} of {len(err)} orbit(s) gated by their own sensitivity")
    for i in bad[:10]:
        print(f"  orbit {i}: err {err[i]:.3e} tol {tol[i]:.3e}")
    assert len(bad) == 0
    assert gated <= max_gated


def test_integrate_orbit_default_nbatch_dense():
    """Default kwargs (atol = rtol = 1e-10, nbatch = 100) on 1,030 orbits:
    ten full batches and a ragged one of 30, dense output at 301 times."""
    from gala_amd.potential import Hamiltonian
    orc = mw_oracle()
    n = 1030
    w0 = _ics(n, 11)
    t = np.linspace(0.0, 600.0, 301)
    ref, st_ref, stats_ref = O.dop853(orc, w0, t, atol=1e-10, rtol=1e-10, nbatch=100,
                                      save_all=True)
    sens = _sensitivity(orc, w0, t, 100, True, atol=1e-10, rtol=1e-10)
    orbit = Hamiltonian(mw_product()).integrate_orbit(w0, t=t, Integrator="dopri853")
    w = orbit.w()
    assert w.shape == (6, 301, n)
    assert np.array_equal(w[:, 0], w0)
    _check(w, ref, sens, "nbatch=100 dense", max_gated=n // 20)
    # the per-orbit answer is not gala's default one: the batches matter
    one = O.dop853(orc, w0, t, atol=1e-10, rtol=1e-10, nbatch=1, save_all=True)[0]
    assert np.max(state_rel_err(one, ref)) > 1e-7


@pytest.mark.parametrize("nbatch,n", [(100, 2050), (37, 500), (64, 640), (300, 700)])
def test_batched_final_states_and_stats(nbatch, n):
    """save_all=False over 1 Gyr: final states, statuses and every batch's
    step counts, for several batch widths (one wave, several waves, the
    1024-thread workgroup kernel) and ragged last batches.  Initial
    conditions as the 1,024-orbit Leapfrog workload's (5-30 kpc, near-circular
    speeds).  Orbits through the 0.07 kpc nucleus are chaotic over Gyr, and a
    batch shares its members' sensitivity through its common step: a batch
    whose orbits are all insensitive in the oracle (one-ulp response below
    1e-10) must take exactly the oracle's steps; the others may not."""
    from gala_amd import _device as dev
    from gala_amd import workloads
    from gala_amd.integrate import dop853_batched
    from tests.helpers import vcirc_oracle
    orc = mw_oracle()
    w0 = np.ascontiguousarray(workloads.composite_leapfrog(vcirc_oracle(orc), n=2050,
                                                           seed=8)[0][:, :n])
    t = np.linspace(0.0, 1000.0, 1001)
    ref, st_ref, stats_ref = O.dop853(orc, w0, t, atol=1e-10, rtol=1e-9, nbatch=nbatch,
                                      save_all=False)
    sens = _sensitivity(orc, w0, t, nbatch, False, atol=1e-10, rtol=1e-9)
    res = dop853_batched(mw_product(), w0, t[0], t[-1], t[1] - t[0], nbatch=nbatch, atol=1e-10,
                         rtol=1e-9)
    st = dev.to_host(res["status"])
    stats = dev.to_host(res["stats"])
    assert np.array_equal(st, st_ref)
    batch = np.arange(n) // nbatch
    nbat = batch[-1] + 1
    calm = np.array([np.all(sens[batch == b] <= 1e-10) for b in range(nbat)])
    same = np.array([np.array_equal(stats[batch == b], stats_ref[batch == b])
                     for b in range(nbat)])
    print(f"nbatch={nbatch}: {same.sum()} of {nbat} batches take the oracle's steps; "
          f"{calm.sum()} batches hold only insensitive orbits")
    assert np.all(same[calm])
    _check(dev.to_host(res["w_final"]), ref, sens, f"nbatch={nbatch} final",
           max_gated=n // 20)


def test_batched_generic_layout_and_hinit():
    """A potential outside the Milky-Way evaluator (generic component loop,
    shifted component), and h0 = 0 (Hairer's initial step over the batch,
    dop853.cpp:18-86)."""
    from gala_amd import _device as dev
    from gala_amd.integrate import dop853_batched
    from gala_amd.potential import HernquistPotential, PlummerPotential
    pot = HernquistPotential(m=1e11, c=2.0) + PlummerPotential(m=1e10, b=1.0,
                                                               origin=[1.0, 0.5, -0.2])
    orc = [O.hernquist(1e11, 2.0), O.plummer(1e10, 1.0, q0=np.array([1.0, 0.5, -0.2]))]
    n = 250
    w0 = _ics(n, 3)
    for h0 in (1.0, 0.0):
        t = np.array([0.0, h0, 800.0]) if h0 else np.array([0.0, 0.0, 800.0])
        ref, st_ref, stats_ref = O.dop853(orc, w0, t, atol=1e-10, rtol=1e-10, nbatch=100,
                                          save_all=False)
        res = dop853_batched(pot, w0, 0.0, 800.0, h0, nbatch=100, atol=1e-10, rtol=1e-10)
        assert np.array_equal(dev.to_host(res["status"]), st_ref)
        err = state_rel_err(dev.to_host(res["w_final"]), ref)
        print(f"generic h0={h0}: max err {err.max():.3e}")
        assert err.max() < 1e-8


def test_batched_failure_is_per_batch():
    """One orbit that cannot finish within nmax fails its whole batch (-2)
    and no other; integrate_orbit raises as gala's err_if_fail does."""
    from gala_amd import _device as dev
    from gala_amd.integrate import dop853_batched
    from gala_amd.potential import Hamiltonian
    orc = mw_oracle()
    n = 250
    w0 = _ics(n, 8)
    w0[:, 130] = [0.05, 0.0, 0.0, 0.0, 0.3, 0.0]  # a tight nuclear orbit: many steps
    t = np.linspace(0.0, 500.0, 501)
    ref, st_ref, _ = O.dop853(orc, w0, t, atol=1e-10, rtol=1e-10, nbatch=100, save_all=False,
                              nmax=2000)
    res = dop853_batched(mw_product(), w0, 0.0, 500.0, 1.0, nbatch=100, nmax=2000)
    st = dev.to_host(res["status"])
    assert np.array_equal(st, st_ref)
    assert np.all(st[100:200] == -2) and np.all(st[:100] == 1) and np.all(st[200:] == 1)
    with pytest.raises(RuntimeError):
        Hamiltonian(mw_product()).integrate_orbit(w0, t=t, Integrator="dopri853",
                                                  Integrator_kwargs=dict(nmax=2000))


@pytest.mark.parametrize("nbatch,n,T,dense,pilot,nmax", [
    (8, 4000, 300.0, False, 1, 0),       # pause after every batch's first attempt
    (8, 4000, 300.0, True, 96, 0),       # dense output across the pause
    (100, 2050, 300.0, True, 40, 0),     # two waves per batch, a ragged last batch
    (37, 1000, 300.0, False, 100000, 0),  # every batch finishes in the first launch
    (100, 1000, 500.0, False, 50, 400),  # a batch failing at nmax after the pause
    (100, 1000, 500.0, False, 1000, 400),  # ... and before it
])
def test_ordered_launches_equal_one_launch(monkeypatch, nbatch, n, T, dense, pilot, nmax):
    """Batches in two launches (a pause after `pilot` attempts, then the
    batches resumed longest predicted remainder first) take exactly the steps
    of one launch in index order: final states, dense output, statuses and
    step counts bit for bit."""
    from gala_amd import _device as dev
    from gala_amd.integrate import dop853_batched
    w0 = _ics(n, 11)
    w0[:, 3] = [0.05, 0.0, 0.0, 0.0, 0.3, 0.0]  # a costly orbit in batch 0
    t_out = np.linspace(0.0, T, 31) if dense else None
    runs = []
    for mode in ("0", "1"):
        monkeypatch.setenv("GAMD_DOP853_BATCH_ORDER", mode)
        monkeypatch.setenv("GAMD_DOP853_BATCH_PILOT", str(pilot))
        r = dop853_batched(mw_product(), w0, 0.0, T, 1.0, nbatch=nbatch, rtol=1e-9, atol=1e-10,
                           nmax=nmax, t_out=t_out)
        runs.append({k: dev.to_host(v) for k, v in r.items()})
    one, two = runs
    st = one["status"]
    print(f"nbatch={nbatch} n={n}: statuses {dict(zip(*np.unique(st, return_counts=True)))}, "
          f"attempts per batch {one['stats'][::nbatch, 0].min()}-{one['stats'][::nbatch, 0].max()}")
    if nmax:
        assert np.any(st == -2)
    for k in one:
        assert np.array_equal(one[k], two[k], equal_nan=True), k


def test_ordered_launches_by_default_for_many_batches(monkeypatch):
    """With more batches than resident workgroups (20,000 orbits in batches
    of 8: 2,500 one-wave workgroups) the two launches are the default; the
    results equal one launch in index order bit for bit."""
    from gala_amd import _device as dev
    from gala_amd.integrate import dop853_batched
    n = 20000
    w0 = _ics(n, 12)
    runs = []
    for mode in ("0", None):
        if mode is None:
            monkeypatch.delenv("GAMD_DOP853_BATCH_ORDER", raising=False)
        else:
            monkeypatch.setenv("GAMD_DOP853_BATCH_ORDER", mode)
        monkeypatch.delenv("GAMD_DOP853_BATCH_PILOT", raising=False)
        r = dop853_batched(mw_product(), w0, 0.0, 200.0, 1.0, nbatch=8, rtol=1e-9, atol=1e-10)
        runs.append({k: dev.to_host(v) for k, v in r.items()})
    for k in runs[0]:
        assert np.array_equal(runs[0][k], runs[1][k]), k


@pytest.mark.parametrize("kind,nbatch,n,pilot,sign", [("generic", 300, 1400, 64, 1.0),
                                                      ("generic", 100, 700, 30, -1.0),
                                                      ("scf", 100, 300, 20, 1.0)])
def test_ordered_launches_other_evaluators_and_wide_batches(monkeypatch, kind, nbatch, n, pilot,
                                                           sign):
    """The two launches with the generic component loop (a shifted
    component), backward in time, with the SCF expansion, and with batches of
    300 (five waves: the 1,024-thread kernel, three waves idle during the
    sums), dense output included: bit for bit the one launch's results."""
    from gala_amd import _device as dev, workloads
    from gala_amd.integrate import dop853_batched
    from gala_amd.potential import HernquistPotential, PlummerPotential, SCFPotential
    if kind == "scf":
        S, Tc = workloads.scf_coefficients(nmax=4, lmax=3, seed=2)
        pot = SCFPotential(m=1e11, r_s=5.0, Snlm=S, Tnlm=Tc)
    else:
        pot = HernquistPotential(m=1e11, c=2.0) + PlummerPotential(m=1e10, b=1.0,
                                                                   origin=[1.0, 0.5, -0.2])
    w0 = _ics(n, 13)
    t_out = sign * np.linspace(0.0, 400.0, 17)
    runs = []
    for mode in ("0", "1"):
        monkeypatch.setenv("GAMD_DOP853_BATCH_ORDER", mode)
        monkeypatch.setenv("GAMD_DOP853_BATCH_PILOT", str(pilot))
        r = dop853_batched(pot, w0, 0.0, sign * 400.0, sign * 1.0, nbatch=nbatch, rtol=1e-9,
                           atol=1e-10, t_out=t_out)
        runs.append({k: dev.to_host(v) for k, v in r.items()})
    assert np.all(runs[0]["status"] == 1)
    assert runs[0]["stats"][:, 0].max() > pilot  # batches resumed in the second launch
    for k in runs[0]:
        assert np.array_equal(runs[0][k], runs[1][k]), k
