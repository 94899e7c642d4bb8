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
