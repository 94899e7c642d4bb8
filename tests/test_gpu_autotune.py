"""The DOP853 launch-shape autotuner: it times every candidate on the
caller's workload, keeps the fastest, and remembers it (memory and JSON
file); the chosen shape gives the same results as any other."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autotune_dop853_picks_and_remembers(tmp_path, monkeypatch):
    import torch
    from gala_amd import _device as dev
    from gala_amd import autotune, workloads
    from gala_amd.integrate import dop853
    from gala_amd.potential import MilkyWayLikePotential
    cache = tmp_path / "tune.json"
    monkeypatch.setenv("GALA_AMD_AUTOTUNE_CACHE", str(cache))
    monkeypatch.setattr(autotune, "_MEM", {})
    pot = MilkyWayLikePotential()
    n = 4096
    w0, T = workloads.composite_dop853(pot.circular_velocity, n=n, seed=3)
    T = T * 0.02
    order = workloads.cost_order(w0, T, pot.gradient)
    w0d, t1d = dev.to_dev(w0), dev.to_dev(T)
    t0d = torch.zeros(n, dtype=torch.float64, device="cuda")
    ordd = dev.to_dev(order, dtype=torch.int32)
    r = autotune.tune_dop853(pot, w0d, t0d, t1d, 1.0, rtol=1e-8, atol=1e-10, order=ordd,
                             blocks=(64, 256), tails=(0, 64))
    assert not r["cached"] and r["block"] in (64, 256) and r["n_tail"] in (0, 64)
    assert len(r["trials_ms"]) == 3 and all(v > 0 for v in r["trials_ms"].values())
    assert autotune.lookup_dop853(pot, n, 1e-8, 1e-10) == (r["block"], r["n_tail"])
    saved = json.loads(cache.read_text())
    assert len(saved) == 1 and list(saved.values())[0]["block"] == r["block"]
    again = autotune.tune_dop853(pot, w0d, t0d, t1d, 1.0, rtol=1e-8, atol=1e-10, order=ordd)
    assert again["cached"] and again["block"] == r["block"]
    # speed knobs only: the tuned shape and the default give identical states
    a = dop853(pot, w0d, t0d, t1d, 1.0, rtol=1e-8, atol=1e-10, order=ordd,
               n_tail=r["n_tail"], block=r["block"])
    b = dop853(pot, w0d, t0d, t1d, 1.0, rtol=1e-8, atol=1e-10)
    assert np.array_equal(dev.to_host(a["w_final"]), dev.to_host(b["w_final"]))


def test_pilot_order_is_a_permutation_and_changes_nothing():
    """order="pilot": a short pilot pass estimates each orbit's step count and
    the queue hands orbits out longest first; the states are those of any
    other order, bit for bit, and the estimate ranks the long orbits first."""
    import torch
    from gala_amd import _device as dev
    from gala_amd import workloads
    from gala_amd.integrate import dop853, pilot_order
    from gala_amd.potential import MilkyWayLikePotential
    pot = MilkyWayLikePotential()
    n = 8192
    w0, T = workloads.composite_dop853(pot.circular_velocity, n=n, seed=5)
    T = T * 0.05
    w0d, t1d = dev.to_dev(w0), dev.to_dev(T)
    t0d = torch.zeros(n, dtype=torch.float64, device="cuda")
    order = dev.to_host(pilot_order(pot, w0d, t0d, t1d, 1.0, rtol=1e-8, atol=1e-10))
    assert order.dtype == np.int32 and np.array_equal(np.sort(order), np.arange(n))
    a = dop853(pot, w0d, t0d, t1d, 1.0, rtol=1e-8, atol=1e-10, order="pilot", n_tail=64)
    b = dop853(pot, w0d, t0d, t1d, 1.0, rtol=1e-8, atol=1e-10)
    assert np.array_equal(dev.to_host(a["w_final"]), dev.to_host(b["w_final"]))
    nstep = dev.to_host(b["stats"])[:, 0]
    top = np.argsort(-nstep)[:64]
    rank = np.empty(n, int)
    rank[order] = np.arange(n)
    # the 64 costliest orbits sit near the head of the pilot's order
    assert np.mean(rank[top] < 512) > 0.8
