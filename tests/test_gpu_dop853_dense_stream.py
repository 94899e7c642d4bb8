"""Streamed DOP853 dense output (gamd_dop853_dense_stream): the t_out grid in
windows, each window's samples copied to pinned host memory while the next
window integrates.  Orbits stop between windows at a step boundary, and a
step that also covers the next window's first sample is finished from its
kept polynomial, so every sample, final state, status and step count must
equal gamd_dop853's single launch with w_out bit for bit -- for every window
size, with and without graph capture, with a queue order, for the Milky-Way,
generic and SCF evaluators, and backward in time.  (gamd_dop853's dense
output itself is held to the oracle in test_gpu_dop853.py.)"""
import numpy as np
import pytest

from gala_amd import workloads
from tests.helpers import mw_oracle, mw_product, vcirc_oracle

pytestmark = pytest.mark.gpu


def _written(t_out, t0, t1):
    """(n_out, N) mask of the samples an orbit reaches (inside [t0, t1])"""
    lo, hi = np.minimum(t0, t1), np.maximum(t0, t1)
    return (t_out[:, None] >= lo[None, :]) & (t_out[:, None] <= hi[None, :])


def _check_equal(direct, streamed, mask, label):
    """samples an orbit writes (inside its [t0, t1], and only for orbits that
    reach t1: one that stops early, e.g. gala's stiffness return -4, writes
    no later samples in either path), final states, statuses, step counts"""
    wd = direct["w_out"].cpu().numpy()
    ws = streamed["w_out"]
    st = direct["status"].cpu().numpy()
    m = np.broadcast_to((mask & (st == 1)[None, :])[None], wd.shape)
    bad = m & ~((wd == ws) | (np.isnan(wd) & np.isnan(ws)))
    if bad.any():
        c, j, i = np.nonzero(bad)
        print(f"{label}: {bad.sum()} samples differ; first (component, sample, orbit): "
              f"{list(zip(c[:5], j[:5], i[:5]))}; orbits {np.unique(i)[:10]}")
    assert not bad.any(), label
    for k in ("w_final", "status", "stats"):
        assert np.array_equal(direct[k].cpu().numpy(), streamed[k].cpu().numpy()), (label, k)


def test_windows_match_one_launch():
    """3,000 orbits of the uneven workload (0.1-10 Gyr each) sampled at 41
    shared times: orbits finish in different windows and, with one sample per
    window, many steps span several windows.  Every lane layout stops at a
    window's end: one lane per orbit, and the 4-lane quads and 16-lane rows
    that take a wave's last orbits and the n_tail longest."""
    from gala_amd.integrate import dop853, dop853_to_host, pilot_order
    pot = mw_product()
    w0, T = workloads.composite_dop853(vcirc_oracle(mw_oracle()), n=3000, seed=31)
    t_out = np.linspace(0.0, 10000.0, 41)
    mask = _written(t_out, np.zeros_like(T), T)
    direct = dop853(pot, w0, 0.0, T, 1.0, rtol=1e-8, atol=1e-10, t_out=t_out)
    st = direct["status"].cpu().numpy()
    print(f"statuses {dict(zip(*np.unique(st, return_counts=True)))}; "
          f"{mask.sum()} samples written of {mask.size}")
    order = pilot_order(pot, w0, 0.0, T, 1.0, rtol=1e-8, atol=1e-10)
    for rows in (1, 7, 41):
        for graph in (False, True):
            r = dop853_to_host(pot, w0, 0.0, T, t_out, h0=1.0, rtol=1e-8, atol=1e-10,
                               window_rows=rows, graph=graph)
            _check_equal(direct, r, mask, f"rows {rows} graph {graph}")
    r = dop853_to_host(pot, w0, 0.0, T, t_out, h0=1.0, rtol=1e-8, atol=1e-10, order=order,
                       window_rows=3, graph=True)
    _check_equal(direct, r, mask, "pilot order")
    # the longest orbits on dedicated 16-lane rows, stopping at each window's end
    for rows, n_tail in ((1, 64), (5, 512), (41, 512)):
        r = dop853_to_host(pot, w0, 0.0, T, t_out, h0=1.0, rtol=1e-8, atol=1e-10, order=order,
                           n_tail=n_tail, window_rows=rows, graph=False)
        _check_equal(direct, r, mask, f"rows {rows} n_tail {n_tail}")


def test_windows_visit_every_orbit_when_slots_outnumber_lanes():
    """131,072 short orbits, more than the persistent grid has lanes, in a
    shuffled queue order with one sample per window: in later windows most
    queue slots name orbits that have finished, and a lane that draws one
    must draw again until the queue is empty, so that no running orbit is
    left out of a window."""
    from gala_amd.integrate import dop853, dop853_to_host
    pot = mw_product()
    w0, T = workloads.composite_dop853(vcirc_oracle(mw_oracle()), n=131072, seed=33)
    T = T * 0.02
    t_out = np.linspace(0.0, 200.0, 21)
    mask = _written(t_out, np.zeros_like(T), T)
    order = np.random.default_rng(4).permutation(T.size).astype(np.int32)
    direct = dop853(pot, w0, 0.0, T, 1.0, rtol=1e-8, atol=1e-10, t_out=t_out)
    for rows, n_tail in ((1, 0), (1, 1024), (4, 256)):
        r = dop853_to_host(pot, w0, 0.0, T, t_out, h0=1.0, rtol=1e-8, atol=1e-10, order=order,
                           n_tail=n_tail, window_rows=rows, graph=False)
        _check_equal(direct, r, mask, f"rows {rows} n_tail {n_tail}")


def test_windows_generic_scf_and_backward():
    from gala_amd.integrate import dop853, dop853_to_host
    from gala_amd.potential import HernquistPotential, PlummerPotential, SCFPotential
    rng = np.random.default_rng(5)
    n = 500
    q = rng.normal(size=(3, n)) * 8.0
    v = rng.normal(size=(3, n)) * 0.15
    w0 = np.vstack([q, v])
    T = rng.uniform(200.0, 3000.0, n)
    t_out = np.linspace(0.0, 3000.0, 25)
    generic = HernquistPotential(m=1e11, c=2.0) + PlummerPotential(m=1e10, b=1.0, origin=[1.0, 0, 0])
    S, Tc = workloads.scf_coefficients(nmax=4, lmax=3, seed=2)
    scf = SCFPotential(m=1e11, r_s=5.0, Snlm=S, Tnlm=Tc)
    for name, pot, sign, k in (("generic", generic, 1.0, n), ("generic backward", generic, -1.0, n),
                               ("scf", scf, 1.0, 64)):
        t1 = sign * T[:k]
        to = sign * t_out
        mask = _written(to, np.zeros(k), t1)
        direct = dop853(pot, w0[:, :k], 0.0, t1, sign * 1.0, rtol=1e-10, atol=1e-10, t_out=to)
        for rows, graph in ((1, False), (4, True)):
            r = dop853_to_host(pot, w0[:, :k], 0.0, t1, to, h0=sign * 1.0, rtol=1e-10,
                               atol=1e-10, window_rows=rows, graph=graph)
            _check_equal(direct, r, mask, f"{name} rows {rows} graph {graph}")


def test_argument_checks():
    """the C-ABI refuses a zero window, a missing host buffer and a short
    workspace before any launch"""
    import torch
    from gala_amd import _device as dev, _lib
    lib = _lib.load()
    pot = mw_product()
    assert lib.gamd_dop853_dense_workspace_bytes(10, 0) == -1
    w0 = dev.to_dev(np.ones((6, 10)))
    t0 = dev.to_dev(np.zeros(10))
    t1 = dev.to_dev(np.ones(10))
    tout = dev.to_dev(np.linspace(0.0, 1.0, 5))
    out = torch.empty((6, 5, 10), dtype=torch.float64, pin_memory=True)
    wf = dev.empty((6, 10))
    need = int(lib.gamd_dop853_dense_workspace_bytes(10, 2))
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")

    def call(rows, host, nbytes):
        return lib.gamd_dop853_dense_stream(pot.c_handle, w0.data_ptr(), 10, t0.data_ptr(),
                                            t1.data_ptr(), None, 1e-8, 1e-8, 0, 0, 0.0,
                                            tout.data_ptr(), 5, rows, 0, host, wf.data_ptr(),
                                            None, None, None, 0, 0, ws.data_ptr(), nbytes,
                                            dev.stream_ptr())
    invalid = -1  # GAMD_ERR_INVALID
    assert call(0, out.data_ptr(), need) == invalid
    assert call(2, None, need) == invalid
    assert call(2, out.data_ptr(), need - 8) == invalid
    assert call(2, out.data_ptr(), need) == 0
    torch.cuda.synchronize()


def test_integrate_orbit_many_orbits_per_orbit_dop853():
    """integrate_orbit with nbatch=1 on 5,000 orbits takes its fast path
    (pilot order, the longest on 16-lane rows, dense output streamed in
    windows); the Orbit it returns equals dop853's dense output and the
    final states of a plain launch, bit for bit."""
    from gala_amd import _device as dev
    from gala_amd.integrate import dop853
    from gala_amd.potential import Hamiltonian
    rng = np.random.default_rng(9)
    n = 5000
    w0 = np.vstack([rng.uniform(3, 30, (3, n)), rng.normal(0, 0.12, (3, n))])
    t = np.linspace(0.0, 800.0, 201)
    kw = dict(atol=1e-10, rtol=1e-10)
    orbit = Hamiltonian(mw_product()).integrate_orbit(
        w0, t=t, Integrator="dopri853", Integrator_kwargs=dict(nbatch=1, **kw))
    ref = dop853(mw_product(), w0, t[0], t[-1], t[1] - t[0], t_out=t, **kw)
    assert np.all(dev.to_host(ref["status"]) == 1)
    assert np.array_equal(orbit.w(), dev.to_host(ref["w_out"]))


def test_autotuner_tries_the_whole_grid_as_one_window(monkeypatch, tmp_path):
    """The window autotuner times the whole t_out grid as one window beside
    its chunk-sized candidates (when that workspace fits), remembers the
    fastest, and dop853_to_host then uses it with the same samples as direct
    writes."""
    from gala_amd import autotune
    from gala_amd.integrate import dop853, dop853_to_host
    monkeypatch.setenv("GALA_AMD_AUTOTUNE_CACHE", str(tmp_path / "autotune.json"))
    pot = mw_product()
    w0, T = workloads.composite_dop853(vcirc_oracle(mw_oracle()), n=2000, seed=32)
    t_out = np.linspace(0.0, 10000.0, 401)
    # 8 and 32 MB of saved rows: windows of 83 and 333 samples, and the whole grid
    tuned = autotune.tune_dop853_dense(pot, w0, 0.0, T, t_out, h0=1.0, rtol=1e-8, atol=1e-10,
                                       mbs=(8, 32), repeats=1, force=True)
    rows = sorted({int(k.split(",")[0].split("=")[1]) for k in tuned["trials_ms"]})
    print(f"window candidates {rows}; picked {tuned['window_rows']} (graph {tuned['graph']})")
    assert rows == [83, 333, 401]
    assert tuned["window_rows"] in rows
    direct = dop853(pot, w0, 0.0, T, 1.0, rtol=1e-8, atol=1e-10, t_out=t_out)
    r = dop853_to_host(pot, w0, 0.0, T, t_out, h0=1.0, rtol=1e-8, atol=1e-10)
    _check_equal(direct, r, _written(t_out, np.zeros_like(T), T), "tuned")
