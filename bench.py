#!/usr/bin/env python
"""Benchmark: orbit-steps/s on one MI355X.

Default workload (the one the metric is quoted on): batched uneven DOP853,
1,048,576 orbits in the Milky-Way-like composite potential, FP64 state,
rtol 1e-8 / atol 1e-10, per-orbit integration time logU(0.1, 10) Gyr, all
orbits in one persistent launch (gamd_dop853).  One "step" of this script is
one full integration of every orbit; orbit-steps are ACCEPTED DOP853 steps
(function evaluations/s reported beside them).  Inputs are resident in HBM
before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints one JSON line with the metric, a roofline object for the dominant
kernel (achieved algorithmic FP64 TFLOP/s from HIP events on the launch
stream against the 78.6 TFLOP/s FP64 peak) and a cpu_baseline object (the
oracle's -Ofast OpenMP restatement of gala's C loops on a bounded sample of
the same workload, on this host's cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "orbit-steps/sec @1 MI355X: exact FP64 vs SCF-surrogate at ≤1e-4 dE/E"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by the copy-kernel calibration), if recorded."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except Exception:
        return None


def host_cpu():
    """Model name, logical CPUs this process may use, and SMT state of the
    host whose cores run the CPU baseline."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    smt = None
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            smt = f.read().strip() == "1"
    except OSError:
        pass
    return {"model": model, "logical_cpus_available": len(os.sched_getaffinity(0)),
            "smt_active": smt}


def cpu_baseline_dop853(w0, T, stride):
    from oracle import oracle as O
    orc = O.milky_way_like()
    sel = np.arange(0, w0.shape[1], stride)
    ws = np.ascontiguousarray(w0[:, sel])
    t0 = time.perf_counter()
    _, st, stats = O.dop853_orbits(orc, ws, 0.0, T[sel], 1.0, atol=1e-10, rtol=1e-8, fast=True)
    dt = time.perf_counter() - t0
    return {"value": float(stats[:, 1].sum() / dt), "unit": "orbit-steps/s",
            "cores": int(O.num_threads(fast=True)), "kind": "port",
            "sample": f"every {stride}th orbit of the same workload ({len(sel)} orbits, "
                      f"{int(stats[:, 1].sum())} accepted steps, {dt:.1f} s), oracle/gala_oracle.c "
                      f"built -Ofast with OpenMP",
            "host": host_cpu()}


def bench_dop853(args):
    import torch
    from gala_amd import _device as dev
    from gala_amd import autotune, roofline, workloads
    from gala_amd.integrate import dop853, pilot_order
    from gala_amd.potential import MilkyWayLikePotential

    n = args.n or 1048576
    pot = MilkyWayLikePotential()
    log(f"generating {n} orbits")
    w0, T = workloads.composite_dop853(pot.circular_velocity, n=n, seed=1)
    w0d = dev.to_dev(w0)
    t0d = torch.zeros(n, dtype=torch.float64, device="cuda")
    t1d = dev.to_dev(T)
    h0d = torch.ones(n, dtype=torch.float64, device="cuda")
    # the queue order of every timed pass comes from a pilot pass inside that
    # pass (order="pilot"); the autotuner's trials use one such order
    ordd = pilot_order(pot, w0d, t0d, t1d, h0d, rtol=1e-8, atol=1e-10)
    torch.cuda.synchronize()
    block, n_tail, tuned = args.block, args.n_tail, None
    if block == 0 or n_tail is None:
        # the project's autotuner: whole passes of this workload per candidate
        # launch shape (results are identical for every candidate)
        tuned = autotune.tune_dop853(pot, w0d, t0d, t1d, h0d, rtol=1e-8, atol=1e-10, order=ordd,
                                     blocks=autotune.BLOCKS if block == 0 else (block,),
                                     tails=autotune.TAILS if n_tail is None else (n_tail,),
                                     repeats=3, log=log)
        block, n_tail = tuned["block"], tuned["n_tail"]
        log(f"autotuned: block {block}, n_tail {n_tail}")

    def one(mid=None):
        # pilot pass (1 % of every orbit's time span) -> longest-first order ->
        # the full pass, all on the device; `mid` marks the full launch's start
        o = pilot_order(pot, w0d, t0d, t1d, h0d, rtol=1e-8, atol=1e-10, block=block)
        if mid is not None:
            mid.record(torch.cuda.current_stream())
        return dop853(pot, w0d, t0d, t1d, h0d, rtol=1e-8, atol=1e-10, order=o,
                      n_tail=n_tail, block=block)

    for i in range(args.warmup):
        one()
        torch.cuda.synchronize()
        log(f"warmup {i + 1}/{args.warmup} done")
    stream = torch.cuda.current_stream()
    evs = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        e0 = torch.cuda.Event(enable_timing=True)
        em = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        res = one(em)
        e1.record(stream)
        evs.append((e0, em, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    # the full launch alone (the roofline's kernel) and the pilot + argsort
    kernel_ms = float(np.mean([m.elapsed_time(b) for a, m, b in evs]))
    pilot_ms = float(np.mean([a.elapsed_time(m) for a, m, b in evs]))
    stats = dev.to_host(res["stats"])
    status = dev.to_host(res["status"])
    naccpt = int(stats[:, 1].sum())
    nrejct = int(stats[:, 2].sum())
    nfcn = int(stats[:, 3].sum())
    ms_per_step = wall * 1000.0 / args.steps
    value = naccpt / (ms_per_step / 1000.0)
    codes, counts = np.unique(status, return_counts=True)
    status_hist = {str(int(c)): int(k) for c, k in zip(codes, counts)}
    log(f"{args.steps} passes: {ms_per_step:.1f} ms/pass, pilot {pilot_ms:.1f} ms + kernel {kernel_ms:.1f} ms; "
        f"{naccpt} accepted, {nrejct} rejected, status counts {status_hist}")
    flops = roofline.dop853_flops(naccpt, nrejct)
    achieved = flops / (kernel_ms / 1000.0) / 1e12
    traffic = traffic_from_profiles("dop853_composite")
    out = {
        "metric": METRIC, "value": value, "unit": "orbit-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded initial conditions, numpy default_rng(1))",
        "metric_note": "value: exact FP64 orbit-steps/s (accepted DOP853 steps) of the "
                       "1,048,576-orbit uneven workload; the exact-against-surrogate half of the "
                       "metric is under scf_exact and scf_surrogate (value, speedup_vs_exact, "
                       "dE_over_E)",
        "config": {"workload": "composite_dop853", "n_orbits": n,
                   "potential": "MW-like composite: Hernquist 1.7e9/0.07 + Hernquist 5e9/1.0 + "
                                "Miyamoto-Nagai 6.8e10/3.0/0.28 + NFW 5.4e11/15.6",
                   "rtol": 1e-8, "atol": 1e-10, "T": "logU(0.1, 10) Gyr per orbit",
                   "r0": "logU(1, 200) kpc", "step": "one integration of every orbit",
                   "orbit_step": "accepted DOP853 step", "parallelism": "single GPU",
                   "queue": f"longest first by a pilot pass over 1 % of each orbit's time "
                            f"(inside the timed pass), {n_tail} on dedicated 16-lane rows",
                   "block": block,
                   "autotune_ms": None if tuned is None else tuned["trials_ms"]},
        "accepted_steps": naccpt, "rejected_steps": nrejct,
        "max_attempts_one_orbit": int(stats[:, 0].max()),
        # gala's codes: 1 success, -4 stiffness detected (dop853.cpp:474-481
        # returns on the first detection; the oracle gives the same orbits -4)
        "status_counts": status_hist,
        "fevals_per_s": nfcn / (ms_per_step / 1000.0),
        "roofline": {"bound": "fp64-valu", "pipe": "fp64 (VALU)", "achieved": achieved,
                     "peak": roofline.PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / roofline.PEAK_FP64_TFLOPS, "traffic": traffic,
                     "kernel": "k_dop853<false,false> (the full launch)",
                     "kernel_ms": kernel_ms, "pilot_ms": pilot_ms,
                     "flops_per_launch": flops},
    }
    return out, (w0, T)


def bench_dop853_batched(args, w0_all):
    """gala's default DOP853 path (Hamiltonian.integrate_orbit passes
    nbatch=100): the first 262,144 orbits of the uneven workload's initial
    conditions over one common 1 Gyr (a batch shares its time grid), final
    states only, nbatch = 100 (gamd_dop853_batched) against nbatch = 1 (the
    per-orbit persistent kernel) on the same orbits.  Accepted steps count
    per orbit (a batch's steps times its orbits)."""
    import torch
    from gala_amd import _device as dev
    from gala_amd import roofline
    from gala_amd.integrate import dop853, dop853_batched
    from gala_amd.potential import MilkyWayLikePotential
    pot = MilkyWayLikePotential()
    n = min(args.n or 262144, w0_all.shape[1], 262144)
    w0d = dev.to_dev(np.ascontiguousarray(w0_all[:, :n]))
    T = 1000.0
    stream = torch.cuda.current_stream()
    res = {}
    for name, fn in (("nbatch100", lambda: dop853_batched(pot, w0d, 0.0, T, 1.0, nbatch=100,
                                                           rtol=1e-8, atol=1e-10)),
                     ("nbatch1", lambda: dop853(pot, w0d, 0.0, T, 1.0, rtol=1e-8, atol=1e-10))):
        fn()
        r, wall, ev = _timed(fn, stream)
        stats = dev.to_host(r["stats"])
        acc = int(stats[:, 1].sum())
        rej = int(stats[:, 2].sum())
        res[name] = {"n_orbits": n, "T": T, "wall_ms": wall * 1e3, "kernel_ms": ev,
                     "accepted_orbit_steps": acc,
                     "value": acc / wall, "unit": "orbit-steps/s",
                     "status_ok": float(np.mean(dev.to_host(r["status"]) == 1))}
        if name == "nbatch100":
            flops = roofline.dop853_flops(acc, rej)
            res[name]["roofline"] = {"bound": "fp64-valu", "pipe": "fp64 (VALU)",
                                     "achieved": flops / (ev / 1e3) / 1e12,
                                     "peak": roofline.PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                                     "frac": flops / (ev / 1e3) / 1e12
                                     / roofline.PEAK_FP64_TFLOPS}
        log(f"dop853 {name}: {n} orbits x {T:.0f} Myr in {wall * 1e3:.1f} ms, "
            f"{acc / wall:.3e} accepted orbit-steps/s")
    return res


def bench_dop853_dense_to_host(args, w0_all, T_all):
    """DOP853 dense output delivered to host memory (what integrate_orbit
    with nbatch=1 returns): the first 262,144 orbits of the uneven workload
    (their own times, up to 10 Gyr) sampled at 101 shared times over 10 Gyr,
    in the pilot's longest-first order with 1,024 orbits on dedicated rows.
    Direct writes (all samples in HBM, then one copy to pinned host memory)
    against gamd_dop853_dense_stream (windows of the time grid, each copied on
    a second stream while the next integrates; window size and graph capture
    autotuned).  Wall clock to a complete host buffer."""
    import torch
    from gala_amd import _device as dev
    from gala_amd import autotune
    from gala_amd.integrate import dop853, dop853_to_host, pilot_order
    from gala_amd.potential import MilkyWayLikePotential
    pot = MilkyWayLikePotential()
    n = min(args.n or 262144, w0_all.shape[1], 262144)
    w0d = dev.to_dev(np.ascontiguousarray(w0_all[:, :n]))
    t1d = dev.to_dev(np.ascontiguousarray(T_all[:n]))
    t_out = np.linspace(0.0, 10000.0, 101)
    kw = dict(rtol=1e-8, atol=1e-10)
    order = pilot_order(pot, w0d, 0.0, t1d, 1.0, **kw)
    host = torch.empty((6, t_out.size, n), dtype=torch.float64, pin_memory=True)

    def direct():
        r = dop853(pot, w0d, 0.0, t1d, 1.0, t_out=t_out, order=order, n_tail=1024, **kw)
        host.copy_(r["w_out"])
        torch.cuda.synchronize()

    tuned = autotune.tune_dop853_dense(pot, w0d, 0.0, t1d, t_out, h0=1.0, order=order,
                                       n_tail=1024, log=log, **kw)

    def streamed():
        dop853_to_host(pot, w0d, 0.0, t1d, t_out, h0=1.0, order=order, n_tail=1024, out=host,
                       **kw)

    direct()
    t_d, t_s = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        direct()
        t_d.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        streamed()
        t_s.append(time.perf_counter() - t0)
    gb = 48.0 * t_out.size * n / 1e9
    log(f"dop853 dense output to host ({n} orbits x {t_out.size} samples, {gb:.2f} GB): "
        f"direct {1e3 * min(t_d):.1f} ms, streamed {1e3 * min(t_s):.1f} ms "
        f"({tuned['window_rows']} samples/window, graph {int(tuned['graph'])})")
    del host
    return {"n_orbits": n, "n_out": int(t_out.size), "GB": gb,
            "direct_ms": 1e3 * min(t_d), "streamed_ms": 1e3 * min(t_s),
            "window_rows": tuned["window_rows"], "graph": tuned["graph"],
            "autotune_ms": tuned["trials_ms"]}


def _energy(pot, w):
    """exact H = Phi + |v|^2/2 of device states w (6, N), on device."""
    from gala_amd import _device as dev
    from gala_amd import _lib
    lib = _lib.load()
    N = w.shape[1]
    out = dev.empty((N,))
    _lib.check(lib.gamd_hamiltonian_energy(pot.c_handle, 0.0, w.contiguous().data_ptr(), N,
                                           out.data_ptr(), dev.stream_ptr()),
               "gamd_hamiltonian_energy")
    return out


def _timed(fn, stream):
    """(result, wall s, HIP-event ms on the launch stream) of one call."""
    import torch
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    res = fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return res, time.perf_counter() - t0, e0.elapsed_time(e1)


def bench_leapfrog(args):
    """Plain FP64 leapfrog in the Milky-Way-like composite: the 1024-orbit,
    10,000-step, dt = 1 Myr case with every step saved (gala's default
    save_all), and 1,048,576 orbits x 1,000 steps keeping the final state
    only and saving every 16th step; the oracle's OpenMP leapfrog on the
    host cores beside each."""
    import torch
    from gala_amd import _device as dev
    from gala_amd import roofline, workloads
    from gala_amd.integrate import leapfrog
    from gala_amd.potential import MilkyWayLikePotential
    from oracle import oracle as O

    pot = MilkyWayLikePotential()
    orc = O.milky_way_like()
    stream = torch.cuda.current_stream()
    out = {}
    w0, dt, n_steps = workloads.composite_leapfrog(pot.circular_velocity, n=1024, seed=0)
    w0d = dev.to_dev(w0)
    leapfrog(pot, w0d, dt, 10)
    _, wall, ev = _timed(lambda: leapfrog(pot, w0d, dt, n_steps, save_stride=1), stream)
    cpu = None
    if not args.no_cpu:
        c0 = time.perf_counter()
        O.leapfrog(orc, w0, np.arange(n_steps + 1) * dt, save_all=True, fast=True)
        cpu = 1024 * n_steps / (time.perf_counter() - c0)
    # 16 waves (one per SIMD of 16 CUs): each step costs one orbit's
    # dependency chain, so neither the FP64 pipe nor HBM bounds it
    flops0 = 1024 * n_steps * roofline.FLOP_LEAPFROG_STEP
    bytes0 = 1024 * (n_steps + 1) * roofline.BYTES_SAVED_STATE
    out["n1024_save_all"] = {"n_orbits": 1024, "n_steps": n_steps, "dt": dt, "save_stride": 1,
                             "value": 1024 * n_steps / wall, "unit": "orbit-steps/s",
                             "kernel_ms": ev, "cpu_value": cpu,
                             "cpu_cores": int(O.num_threads(fast=True)),
                             "gpu_over_cpu": (1024 * n_steps / wall) / cpu if cpu else None,
                             "roofline": {"bound": "latency", "pipe": "fp64 (VALU)",
                                          "achieved": flops0 / (ev / 1e3) / 1e12,
                                          "peak": roofline.PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                                          "frac": flops0 / (ev / 1e3) / 1e12
                                          / roofline.PEAK_FP64_TFLOPS,
                                          "hbm_write_GBps": bytes0 / (ev / 1e3) / 1e9,
                                          "hbm_frac": bytes0 / (ev / 1e3) / 1e9
                                          / roofline.PEAK_HBM_GBS}}
    n = args.lf_n or 1048576
    steps = 1000
    w0, _, _ = workloads.composite_leapfrog(pot.circular_velocity, n=n, seed=0)
    w0d = dev.to_dev(w0)
    leapfrog(pot, w0d, dt, 10, save=False)
    _, wall_f, ev_f = _timed(lambda: leapfrog(pot, w0d, dt, steps, save=False), stream)
    _, wall_s, ev_s = _timed(lambda: leapfrog(pot, w0d, dt, steps, save_stride=16), stream)
    ns = min(n, 65536)
    cpu = None
    if not args.no_cpu:
        c0 = time.perf_counter()
        O.leapfrog(orc, np.ascontiguousarray(w0[:, :ns]), np.arange(steps + 1) * dt,
                   save_all=False, fast=True)
        cpu = ns * steps / (time.perf_counter() - c0)
    flops = n * steps * roofline.FLOP_LEAPFROG_STEP
    saved = 6 * (steps // 16 + 1) * n * 8
    out["n1M_final"] = {
        "n_orbits": n, "n_steps": steps, "dt": dt, "value": n * steps / wall_f,
        "unit": "orbit-steps/s", "kernel_ms": ev_f,
        "roofline": {"bound": "fp64-valu", "pipe": "fp64 (VALU)", "achieved": flops / (ev_f / 1e3) / 1e12,
                     "peak": roofline.PEAK_FP64_TFLOPS, "unit": "TFLOP/s"},
        "cpu_value": cpu, "cpu_cores": int(O.num_threads(fast=True)),
        "cpu_sample": f"first {ns} orbits, {steps} steps, final state only"}
    out["n1M_final"]["roofline"]["frac"] = out["n1M_final"]["roofline"]["achieved"] / \
        roofline.PEAK_FP64_TFLOPS
    out["n1M_final"]["gpu_over_cpu"] = out["n1M_final"]["value"] / cpu if cpu else None
    out["n1M_stride16"] = {"n_orbits": n, "n_steps": steps, "save_stride": 16,
                           "value": n * steps / wall_s, "unit": "orbit-steps/s", "kernel_ms": ev_s,
                           "hbm_write_GBps": saved / (ev_s / 1e3) / 1e9}
    log(f"leapfrog composite: 1024x{n_steps} save_all {out['n1024_save_all']['value']:.3e}/s, "
        f"{n}x{steps} {out['n1M_final']['value']:.3e}/s (CPU {cpu}/s)")
    out["to_host"] = bench_leapfrog_to_host(args, pot, log)
    return out


def bench_leapfrog_to_host(args, pot, log):
    """Saved trajectories delivered to host memory (what integrate_orbit
    returns): direct writes (the whole trajectory in HBM, then one copy to
    pinned host memory) against gamd_leapfrog_stream (chunks copied on a
    second stream while the next chunk computes; chunk size and graph capture
    autotuned), for the 1,024-orbit save_all case and 1M orbits saved every
    16th step.  Wall clock to a complete host buffer."""
    import torch
    from gala_amd import _device as dev
    from gala_amd import autotune, workloads
    from gala_amd.integrate import leapfrog, leapfrog_to_host
    res = {}
    for name, n, steps, stride in (("n1024_save_all", 1024, 10000, 1),
                                   ("n1M_stride16", args.lf_n or 1048576, 1000, 16)):
        w0, dt, _ = workloads.composite_leapfrog(pot.circular_velocity, n=n, seed=0)
        w0d = dev.to_dev(w0)
        nsave = steps // stride + 1
        host = torch.empty((6, nsave, n), dtype=torch.float64, pin_memory=True)

        def direct():
            wout, _ = leapfrog(pot, w0d, dt, steps, save_stride=stride)
            host.copy_(wout)
            torch.cuda.synchronize()

        tuned = autotune.tune_leapfrog_stream(pot, w0d, dt, steps, stride, log=log)

        def streamed():
            leapfrog_to_host(pot, w0d, dt, steps, stride, out=host)

        direct()
        t_d, t_s = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            direct()
            t_d.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            streamed()
            t_s.append(time.perf_counter() - t0)
        gb = 48.0 * nsave * n / 1e9
        res[name] = {"n_orbits": n, "n_steps": steps, "save_stride": stride, "GB": gb,
                     "direct_ms": 1e3 * min(t_d), "streamed_ms": 1e3 * min(t_s),
                     "streamed_GBps": gb / min(t_s), "chunk_rows": tuned["chunk_rows"],
                     "graph": tuned["graph"], "autotune_ms": tuned["trials_ms"]}
        log(f"leapfrog to host ({name}, {gb:.2f} GB): direct {1e3 * min(t_d):.1f} ms, "
            f"streamed {1e3 * min(t_s):.1f} ms ({tuned['chunk_rows']} rows/chunk, graph "
            f"{int(tuned['graph'])})")
        del host
    return res


def bench_mockstream(args):
    """The mock stream of a Plummer progenitor (1e5 Msun, 10 pc) at (15, 0, 0)
    kpc, (0, 180, 30) km/s in the composite: 4000 steps of dt = -1 Myr back,
    2 particles per step (8002), each integrated forward to t = 0 with the
    progenitor's gravity; gala's default DOP853 generator and the Leapfrog
    one, end to end through MockStreamGenerator.run (host DF sampling
    included; the median of five calls after a warm-up), with the oracle's
    OpenMP release loop on the host beside the Leapfrog kernel."""
    import torch
    from gala_amd.dynamics import FardalStreamDF, MockStreamGenerator
    from gala_amd.potential import MilkyWayLikePotential, PlummerPotential
    from gala_amd import workloads
    prog_w0, prog_mass, b, dt, nsteps = workloads.mockstream_progenitor()
    pot = MilkyWayLikePotential()
    out = {}
    from gala_amd import roofline
    for integ in ("leapfrog", "dop853"):
        gens = []

        def run():
            df = FardalStreamDF(gala_modified=True, random_state=np.random.RandomState(1))
            gen = MockStreamGenerator(df, pot,
                                      progenitor_potential=PlummerPotential(m=prog_mass, b=b))
            gen.record_kernel_time = True
            gens.append(gen)
            return gen.run(prog_w0, prog_mass, dt=dt, n_steps=nsteps, Integrator=integ)
        run()
        walls = []
        for _ in range(5):  # one call is milliseconds: the median of five
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            stream, _ = run()
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - t0)
        wall = float(np.median(walls))
        n = stream.xyz.shape[1]
        # particle-steps of the reference's fixed-step loop: release i runs
        # nsteps - i steps for each of its particles; each release's
        # progenitor copy steps alongside (mockstream.pyx:584-590)
        psteps = int(2 * sum(range(nsteps + 1)))
        out[integ] = {"particles": n, "wall_s": wall, "wall_s_min": min(walls),
                      "wall_s_max": max(walls),
                      "particle_steps_per_s": psteps / wall if integ == "leapfrog" else None}
        ev = gens[-1].kernel_events
        if integ == "leapfrog" and ev is not None:
            kms = ev[0].elapsed_time(ev[1])
            flops = psteps * roofline.FLOP_STREAM_PARTICLE_STEP + \
                (psteps // 2) * roofline.FLOP_LEAPFROG_STEP
            out[integ]["kernel_ms"] = kms
            out[integ]["kernel_particle_steps_per_s"] = psteps / (kms / 1e3)
            # ~125 waves for 8,002 particles: one orbit's step chain per wave
            # bounds the kernel, not the FP64 pipe
            out[integ]["roofline"] = {
                "bound": "latency", "pipe": "fp64 (VALU)",
                "kernel": "k_mockstream_pair (release-and-integrate, progenitor copy and particle on a lane pair)",
                "achieved": flops / (kms / 1e3) / 1e12, "peak": roofline.PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": flops / (kms / 1e3) / 1e12 / roofline.PEAK_FP64_TFLOPS}
        log(f"mock stream ({integ}): {n} particles, {wall * 1e3:.1f} ms end to end (median of "
            f"5; {min(walls) * 1e3:.1f}-{max(walls) * 1e3:.1f})")
    if not args.no_cpu:
        from oracle import oracle as O
        rel_t = -np.arange(nsteps + 1)[::-1] * 1.0
        full_t = rel_t
        w0 = np.tile(prog_w0, (2 * (nsteps + 1), 1))
        c0 = time.perf_counter()
        O.mockstream_leapfrog(O.milky_way_like(), [O.plummer(prog_mass, b)], full_t, prog_w0,
                              rel_t, np.full(nsteps + 1, 2, dtype=np.int32), w0, 0.0, fast=True)
        cpu = time.perf_counter() - c0
        out["leapfrog"]["cpu_release_loop_s"] = cpu
        out["leapfrog"]["cpu_cores"] = int(O.num_threads(fast=True))
    return out


def bench_scf(args):
    """Exact SCF leapfrog (nmax 10, lmax 6, m 1e12, r_s 20) and the learned
    surrogate on the same 262,144 orbits x 10,000 steps of dt = 0.5 Myr."""
    import torch
    from gala_amd import _device as dev
    from gala_amd import roofline, workloads
    from gala_amd.integrate import leapfrog
    from gala_amd.potential import SCFPotential
    from gala_amd.surrogate import SCFSurrogate

    S, T = workloads.scf_coefficients()
    pot = SCFPotential(m=1e12, r_s=20.0, Snlm=S, Tnlm=T)
    n = args.scf_n or 262144
    w0, dt, n_steps = workloads.scf_leapfrog(pot.circular_velocity, n=n, seed=3)
    n_steps = args.scf_steps or n_steps
    w0d = dev.to_dev(w0)
    stream = torch.cuda.current_stream()
    E0 = _energy(pot, w0d)

    def rel_drift(w):
        return dev.to_host(((_energy(pot, w) - E0) / E0).abs())

    leapfrog(pot, w0d, dt, 10, save=False)  # warm-up
    (_, wf), wall, ev = _timed(lambda: leapfrog(pot, w0d, dt, n_steps, save=False), stream)
    d_exact = rel_drift(wf)
    exact = {"workload": "scf_leapfrog_exact", "n_orbits": n, "n_steps": n_steps, "dt": dt,
             "value": n * n_steps / wall, "unit": "orbit-steps/s", "wall_s": wall,
             "kernel_ms": ev, "dE_over_E": {"median": float(np.median(d_exact)),
                                            "p99": float(np.percentile(d_exact, 99))},
             "roofline": {"bound": "fp64-valu", "pipe": "fp64 (VALU)",
                          "achieved": n * n_steps * roofline.FLOP_SCF_LEAPFROG_STEP / (ev / 1e3) / 1e12,
                          "peak": roofline.PEAK_FP64_TFLOPS, "unit": "TFLOP/s"}}
    exact["roofline"]["frac"] = exact["roofline"]["achieved"] / roofline.PEAK_FP64_TFLOPS
    log(f"exact SCF leapfrog: {exact['value']:.3e} orbit-steps/s")
    if not args.no_cpu:
        from oracle import oracle as O
        ns, nst = min(n, 4096), 200
        orc = [O.scf(1e12, 20.0, S, T)]
        c0 = time.perf_counter()
        O.leapfrog(orc, np.ascontiguousarray(w0[:, :ns]), np.arange(nst + 1) * dt,
                   save_all=False, fast=True)
        cdt = time.perf_counter() - c0
        exact["cpu_baseline"] = {
            "value": ns * nst / cdt, "unit": "orbit-steps/s", "cores": int(O.num_threads(fast=True)),
            "kind": "port", "sample": f"first {ns} orbits of the same workload, {nst} steps, final "
                                      f"state only ({cdt:.1f} s), oracle/gala_oracle.c built "
                                      f"-Ofast with OpenMP (gala's SCF series with GSL's special "
                                      f"functions replaced by recurrences)",
            "host": host_cpu()}
        exact["gpu_over_cpu"] = exact["value"] / exact["cpu_baseline"]["value"]
        log(f"exact SCF leapfrog CPU: {exact['cpu_baseline']['value']:.3e} orbit-steps/s")

    sur = SCFSurrogate(pot, seed=0)
    (hist), t_train, _ = _timed(lambda: sur.train(n_points=args.train_points, batch=32768,
                                                  epochs=args.train_epochs, lr=2e-3,
                                                  lr_min=2e-5, seed=1, log=log), stream)
    qs = sur.sample_points(1 << 20, seed=99)
    _, g_net = sur.forward(qs)
    g_ex = pot.gradient(qs)
    e_f = np.linalg.norm(dev.to_host(g_net) - g_ex, axis=0) / np.linalg.norm(g_ex, axis=0)
    sur.leapfrog(w0d, dt, 2, tol=1.0)  # warm-up
    res, wall_s, ev_s = _timed(lambda: sur.leapfrog(w0d, dt, n_steps, tol=args.tol), stream)
    fb = dev.to_host(res["fallback_step"])
    d_sur = rel_drift(res["w_final"])
    on = fb < 0
    n_short = max(1, n_steps // 10)
    res2, wall_p, ev_p = _timed(lambda: sur.leapfrog(w0d, dt, n_short, tol=1e300), stream)
    d_pure = rel_drift(res2["w_final"])
    _, wall_e2, _ = _timed(lambda: leapfrog(pot, w0d, dt, n_short, save=False), stream)
    surrogate = {
        "workload": "scf_leapfrog_surrogate", "network": "4 -> 4 x 256 SiLU -> 4, BF16 MFMA",
        "train_points": args.train_points, "train_epochs": args.train_epochs,
        "train_s": t_train, "train_final_mse": float(hist[-1]),
        "force_rel_err": {"median": float(np.median(e_f)), "p90": float(np.percentile(e_f, 90)),
                          "p99": float(np.percentile(e_f, 99)), "points": int(qs.shape[1])},
        "fallback_tol": args.tol, "value": n * n_steps / wall_s, "unit": "orbit-steps/s",
        "wall_s": wall_s, "fraction_fallen_back": float(np.mean(fb >= 0)),
        "median_switch_step": float(np.median(fb[~on])) if np.any(~on) else None,
        # energy errors against the FP64 SCF (the monitor bounds the
        # network's own energy Phi_net + |v|^2/2, not this one)
        "dE_over_E": {"median": float(np.median(d_sur)), "p99": float(np.percentile(d_sur, 99))},
        "dE_over_E_kept_on_network": {
            "median": float(np.median(d_sur[on])) if np.any(on) else None,
            "p90": float(np.percentile(d_sur[on], 90)) if np.any(on) else None,
            "p99": float(np.percentile(d_sur[on], 99)) if np.any(on) else None},
        "note": "the 4x256 network costs ~190 times the FP64 nmax=10, lmax=6 expansion per "
                "evaluation (its conservative force needs a forward and a reverse pass); at "
                "the 32:1 BF16:FP64 peak ratio of the MI355X it cannot be faster than the "
                "exact kernel, let alone 5x",
        "surrogate_only": {"n_steps": n_short, "value": n * n_short / wall_p,
                           "exact_value_same_steps": n * n_short / wall_e2,
                           "dE_over_E": {"median": float(np.median(d_pure)),
                                         "p99": float(np.percentile(d_pure, 99))},
                           "roofline": {"bound": "mfma", "pipe": "bf16 MFMA",
                                        "achieved": n * n_short * roofline.FLOP_MLP_FORCE
                                        / (ev_p / 1e3) / 1e12,
                                        "peak": roofline.PEAK_BF16_TFLOPS, "unit": "TFLOP/s"}},
        "speedup_vs_exact": (n * n_steps / wall_s) / exact["value"],
    }
    rl = surrogate["surrogate_only"]["roofline"]
    rl["frac"] = rl["achieved"] / rl["peak"]
    log(f"surrogate: {surrogate['value']:.3e} orbit-steps/s, fallback "
        f"{surrogate['fraction_fallen_back']:.3f}, force err median {np.median(e_f):.2e}")
    return exact, surrogate


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="composite_dop853")
    ap.add_argument("--n", type=int, default=0, help="override the orbit count (testing)")
    ap.add_argument("--block", type=int, default=0, help="DOP853 workgroup size (0: autotune)")
    ap.add_argument("--n-tail", type=int, default=None,
                    help="longest orbits integrated on dedicated 16-lane rows from the start")
    ap.add_argument("--cpu-stride", type=int, default=4,
                    help="CPU baseline sample: every k-th orbit")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-scf", action="store_true",
                    help="skip the exact-SCF and surrogate measurements")
    ap.add_argument("--no-leapfrog", action="store_true",
                    help="skip the composite leapfrog measurements")
    ap.add_argument("--lf-n", type=int, default=0, help="override the large leapfrog orbit count")
    ap.add_argument("--no-stream", action="store_true", help="skip the mock-stream measurements")
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the DOP853 dense-output-to-host measurement")
    ap.add_argument("--scf-n", type=int, default=0, help="override the SCF orbit count")
    ap.add_argument("--scf-steps", type=int, default=0, help="override the SCF step count")
    ap.add_argument("--train-points", type=int, default=16 * 1024 * 1024)
    ap.add_argument("--train-epochs", type=int, default=8)
    ap.add_argument("--tol", type=float, default=1e-5,
                    help="surrogate energy-drift bound for the exact-SCF fallback")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.exit("bench.py measures one GPU only: gala_amd runs on one GPU and has no "
                 "multi-GPU path (--gpus must be 1)")
    import torch
    if not torch.cuda.is_available():
        sys.exit("bench.py needs a GPU")
    if args.workload != "composite_dop853":
        sys.exit(f"unknown workload {args.workload}")
    out, (w0, T) = bench_dop853(args)
    if not args.no_leapfrog:
        out["leapfrog_composite"] = bench_leapfrog(args)
    out["dop853_default_batching"] = bench_dop853_batched(args, w0)
    if not args.no_dense:
        out["dop853_dense_to_host"] = bench_dop853_dense_to_host(args, w0, T)
    if not args.no_stream:
        out["mockstream"] = bench_mockstream(args)
    if not args.no_scf:
        out["scf_exact"], out["scf_surrogate"] = bench_scf(args)
    if not args.no_cpu:
        log("timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline_dop853(w0, T, args.cpu_stride)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
