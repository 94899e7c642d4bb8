"""Host-side logic (no GPU): the C ABI library exports what the header
declares, the gala-shaped API's argument handling, time grids, integrator
lookup, parameter packing and the seeded workloads."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "gala_amd.h")).read()
    return sorted(set(re.findall(r"\b(gamd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from gala_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libgala_amd.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) <= set(_lib.SIGNATURES), set(syms) - set(_lib.SIGNATURES)
    assert _lib.load().gamd_version().startswith(b"gala_amd")


def test_invalid_arguments_are_rejected_without_gpu_work():
    """Argument checks happen before any HIP call."""
    from gala_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libgala_amd.so not built")
    lib = _lib.load()
    out = ctypes.c_void_p()
    assert lib.gamd_potential_create(None, 9, ctypes.byref(out)) == -1
    assert lib.gamd_potential_create(None, -1, ctypes.byref(out)) == -1
    assert lib.gamd_leapfrog(None, None, 1, 0.0, 1.0, 1, None, 1, None, None, None) == -1
    assert lib.gamd_dop853(None, None, 1, None, None, None, 1e-8, 1e-10, 0, 0, 0.0, None, 0, None,
                           None, None, None, None, 0, 0, None) == -1
    assert lib.gamd_ruth4(None, None, 1, 0.0, 1.0, 1, 1, None, None, None) == -1
    assert lib.gamd_mockstream_dop853(None, None, 1, None, 1, None, None, None, None, 0.0, 1.0,
                                      1e-10, 1e-10, 0, None, None, None, None, None) == -1
    assert lib.gamd_mockstream_leapfrog(None, None, 1, None, 0, 0.0, 1.0, None, None, 0, 0.0,
                                        None, None) == -1
    assert lib.gamd_nbody_leapfrog(None, None, 0, None, 1, 0.0, 1.0, 1, 1, None, None,
                                   None) == -1
    assert lib.gamd_nbody_ruth4(None, None, 0, None, 1, 0.0, 1.0, 1, 1, None, None, None) == -1
    assert lib.gamd_nbody_dop853(None, None, 0, None, 1, 0.0, 1.0, 1.0, 1e-10, 1e-10, 0, -1,
                                 0.0, None, 0, None, None, None, None, None) == -1
    # more massive bodies than the kernels carry, or more bodies than a wave holds
    many = _lib.body_array([None] * (_lib.MAX_MASSIVE + 1))
    assert lib.gamd_nbody_leapfrog(None, many, _lib.MAX_MASSIVE + 1, None, 16, 0.0, 1.0, 1, 1,
                                   None, None, None) == -1
    assert lib.gamd_nbody_dop853(None, many, _lib.MAX_MASSIVE + 1, None, 16, 0.0, 1.0, 1.0,
                                 1e-10, 1e-10, 0, -1, 0.0, None, 0, None, None, None, None,
                                 None) == -1
    assert lib.gamd_potential_destroy(None) == 0
    # the dense-output stream: null arguments, a zero window, a short workspace
    assert lib.gamd_dop853_dense_workspace_bytes(10, 0) == -1
    assert lib.gamd_dop853_dense_workspace_bytes(10, 3) == 256 + (2 * 6 * 3 + 79) * 10 * 8
    assert lib.gamd_dop853_dense_stream(None, None, 1, None, None, None, 1e-8, 1e-10, 0, 0, 0.0,
                                        None, 0, 1, 0, None, None, None, None, None, 0, 0, None,
                                        0, None) == -1


def test_batched_dop853_arguments_rejected():
    """gamd_dop853_batched rejects null pointers and a batch wider than one
    workgroup before any HIP call."""
    from gala_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libgala_amd.so not built")
    lib = _lib.load()
    args = [0.0, 1.0, 1.0, 1e-10, 1e-10, 0, 0, 0.0]
    assert lib.gamd_dop853_batched(None, None, 1, *args, 100, None, 0, None, None, None, None,
                                   None) == -1
    fake = ctypes.c_void_p(8)  # never dereferenced: the nbatch checks come first
    assert lib.gamd_dop853_batched(fake, fake, 1, *args, 0, None, 0, None, fake, None, None,
                                   None) == -1
    assert lib.gamd_dop853_batched(fake, fake, 1, *args, 1025, None, 0, None, fake, None, None,
                                   None) == -4


def test_device_argument_checks_run_on_the_host():
    """Shapes, per-orbit lengths and queue orders are checked before any
    device work: a (5, N) w0, a t1 of the wrong length or an order that is not
    a permutation raise ValueError (they would otherwise be read past their
    ends or leave orbits unvisited)."""
    import torch
    from gala_amd import _device as dev
    assert dev.check_w0(np.zeros((6, 7))) == 7
    assert dev.check_w0(torch.zeros(6, 3, dtype=torch.float32)) == 3
    for bad in (np.zeros((5, 7)), np.zeros(6), np.zeros((6, 2, 2))):
        with pytest.raises(ValueError):
            dev.check_w0(bad)
    dev.check_per_orbit(1.0, 7, "t1")
    dev.check_per_orbit(np.ones(7), 7, "t1")
    with pytest.raises(ValueError):
        dev.check_per_orbit(np.ones(6), 7, "t1")
    dev.check_permutation(np.array([2, 0, 1]), 3)
    dev.check_permutation(torch.tensor([1, 0], dtype=torch.int32), 2)
    for bad in (np.array([0, 1]), np.array([0, 0, 1]), np.array([0, 1, 3]),
                np.array([-1, 0, 1]), torch.tensor([0, 2, 2])):
        with pytest.raises(ValueError):
            dev.check_permutation(bad, 3)
    # the entry points check before they need a GPU
    from gala_amd.integrate import dop853, dop853_to_host
    t_out = np.linspace(0.0, 1.0, 5)
    for fn, extra in ((dop853, {}), (dop853_to_host, {"t_out": t_out})):
        with pytest.raises(ValueError):
            fn(None, np.zeros((5, 4)), 0.0, 1.0, **extra)
        with pytest.raises(ValueError):
            fn(None, np.zeros((6, 4)), 0.0, np.ones(3), **extra)
        with pytest.raises(ValueError):
            fn(None, np.zeros((6, 3)), 0.0, 1.0, order=np.array([0, 0, 1]), **extra)


def test_parse_time_specification():
    from gala_amd.integrate import parse_time_specification as p
    t = p(dt=1.0, n_steps=10)
    assert len(t) == 11 and t[0] == 0.0 and t[-1] == 10.0
    t = p(dt=-1.0, n_steps=4000)
    assert len(t) == 4001 and t[-1] == -4000.0
    t = p(dt=0.1, n_steps=5, t1=2.0)
    np.testing.assert_allclose(t, 2.0 + np.cumsum(np.append([0.0], np.full(5, 0.1))))
    t = p(dt=0.5, t1=0.0, t2=2.0)
    np.testing.assert_array_equal(t, [0.0, 0.5, 1.0, 1.5])
    t = p(dt=-0.5, t1=2.0, t2=0.2)
    assert t[0] == 2.0 and t[-1] == 0.2
    t = p(n_steps=5, t1=0.0, t2=1.0)
    np.testing.assert_array_equal(t, np.linspace(0, 1, 5))
    arr = np.array([0.0, 0.3, 1.0])
    np.testing.assert_array_equal(p(t=arr), arr)
    with pytest.raises(ValueError):
        p(dt=1.0, t1=5.0, t2=0.0)
    with pytest.raises(ValueError):
        p()


def test_get_integrator():
    from gala_amd.integrate import DOPRI853Integrator, LeapfrogIntegrator, get_integrator
    assert get_integrator("Leapfrog") is LeapfrogIntegrator
    assert get_integrator("dop853") is DOPRI853Integrator
    assert get_integrator("DOPRI853") is DOPRI853Integrator
    assert get_integrator(LeapfrogIntegrator) is LeapfrogIntegrator
    with pytest.raises(ValueError):
        get_integrator("euler")


def test_parameter_packing():
    from gala_amd import _lib
    from gala_amd.potential import (CompositePotential, HernquistPotential,
                                    MilkyWayLikePotential, MiyamotoNagaiPotential, NFWPotential,
                                    PlummerPotential, SCFPotential)
    from gala_amd.units import G
    h = HernquistPotential(5e9, 1.0)
    c = h._components()[0]
    assert c["type"] == _lib.HERNQUIST and c["pars"] == [G, 5e9, 1.0] and c["shift_rotate"] == 0
    assert HernquistPotential(m=5e9, c=1.0).parameters == h.parameters
    mn = MiyamotoNagaiPotential(m=6.8e10, a=3.0, b=0.28, origin=[1.0, 0, 0])
    c = mn._components()[0]
    assert c["pars"] == [G, 6.8e10, 3.0, 0.28] and c["shift_rotate"] == 1
    nfw = NFWPotential(m=5.4e11, r_s=15.6)
    assert nfw._components()[0]["pars"] == [G, 5.4e11, 15.6, 1.0, 1.0, 1.0]
    with pytest.raises(NotImplementedError):
        NFWPotential(m=1e12, r_s=20.0, c=0.8)
    with pytest.raises(TypeError):
        PlummerPotential(m=1.0)
    mw = MilkyWayLikePotential()
    types = [c["type"] for c in mw._components()]
    assert types == [_lib.HERNQUIST, _lib.HERNQUIST, _lib.MIYAMOTONAGAI, _lib.NFW]
    comp = CompositePotential()
    comp["a"] = h
    comp["b"] = mn
    assert len(comp._components()) == 2
    both = h + mn
    assert isinstance(both, CompositePotential) and len(both._components()) == 2
    S = np.zeros((3, 2, 2))
    S[0, 0, 0] = 1.0
    scf = SCFPotential(m=1e12, r_s=20.0, Snlm=S)
    c = scf._components()[0]
    assert c["type"] == _lib.SCF and c["S"].shape == (3, 2, 2) and np.all(c["T"] == 0)
    arr, keep = _lib.make_components(mw._components())
    assert arr[3].type == _lib.NFW and arr[3].pars[2] == 15.6


def test_workloads_are_seeded_and_shaped():
    from gala_amd import workloads

    def vc(q):
        return np.full(q.shape[1], 0.2)
    w0a, dt, n = workloads.composite_leapfrog(vc, n=64)
    w0b, _, _ = workloads.composite_leapfrog(vc, n=64)
    assert w0a.shape == (6, 64) and dt == 1.0 and n == 10000
    np.testing.assert_array_equal(w0a, w0b)
    r = np.linalg.norm(w0a[:3], axis=0)
    assert np.all((r >= 5) & (r <= 30))
    # velocity is tangential apart from the 10 km/s noise
    vr = np.sum(w0a[:3] * w0a[3:], axis=0) / r
    assert np.max(np.abs(vr)) < 5 * 10 * 1.0227e-3
    w0, T = workloads.composite_dop853(vc, n=1000)
    assert np.all((T >= 100) & (T <= 10000))
    r = np.linalg.norm(w0[:3], axis=0)
    assert np.all((r >= 1) & (r <= 200))
    S, Tc = workloads.scf_coefficients()
    assert S.shape == (11, 7, 7) and S[0, 0, 0] == 1.0 and np.all(Tc[:, :, 0] == 0)
    for l in range(7):
        for m in range(l + 1, 7):
            assert np.all(S[:, l, m] == 0) and np.all(Tc[:, l, m] == 0)
    order = workloads.cost_order(w0, T, lambda q: -0.01 * q)
    assert sorted(order) == list(range(1000))


def test_units():
    from gala_amd import units
    assert abs(units.G - 4.498502151469553e-12) < 1e-26
    assert abs(units.KMS_TO_KPCMYR - 1.0227121650456959e-3) < 1e-15


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


def test_no_cpu_fallback_in_product():
    """The product never imports the oracle and has no CPU path."""
    pkg = os.path.join(ROOT, "gala_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower() or f == "__init__.py", f


def test_release_counts_matches_isclose_loop():
    """The vectorised release bookkeeping equals the reference's loop
    (mockstream_generator.py:272-276) on grids with exact, near-tolerance,
    overlapping and missing release times."""
    from gala_amd.dynamics import release_counts
    rng = np.random.default_rng(0)

    def loop(orbit_t, t1s, counts):
        out = np.zeros(len(orbit_t), dtype=int)
        for t1, n in zip(t1s, counts):
            out[np.isclose(orbit_t, t1)] = n
        return out
    grids = [np.arange(4001) * 1.0 - 4000.0, np.sort(rng.uniform(-100, 100, 500)),
             np.array([0.0, 1e-9, 2e-8, 1.0, 1.0 + 1e-6, 1.00002, 5.0]),
             np.linspace(0, 1, 11)[::-1]]
    for g in grids:
        for t1s in (np.unique(g), np.unique(g[::3]),
                    np.unique(np.concatenate([g[::2] + rng.normal(0, 1e-5, len(g[::2])),
                                              g[:3] + 2e-8]))):
            counts = rng.integers(1, 5, len(t1s))
            assert np.array_equal(release_counts(g, t1s, counts), loop(g, t1s, counts))
