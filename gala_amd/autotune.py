"""Launch-shape autotuner: the persistent DOP853 kernel's workgroup size and
dedicated rows, the streamed leapfrog output's chunk size and the streamed
DOP853 dense output's window size (with or without graph capture).

gamd_dop853 gives identical results for every workgroup size and every
number of dedicated-row orbits (n_tail), so these are pure speed knobs.  The
tuner times whole passes of the caller's own workload for each candidate
(one knob at a time: the workgroup size first, with the default n_tail when
it is a candidate, then n_tail with the best size), keeps the fastest, and remembers it per device,
potential, tolerance and problem size, in memory and in a JSON file
(``$GALA_AMD_AUTOTUNE_CACHE``, default ``~/.cache/gala_amd/autotune.json``;
set it to an empty string to keep results in memory only).
"""
import json
import os

from . import _device as dev

BLOCKS = (64, 128, 192, 256)  # whole waves up to the kernel's launch bound
TAILS = (0, 256, 1024, 4096)
DEFAULT_BLOCK, DEFAULT_TAIL = 256, 1024

_MEM = {}


def _cache_path():
    p = os.environ.get("GALA_AMD_AUTOTUNE_CACHE")
    if p is None:
        p = os.path.join(os.path.expanduser("~"), ".cache", "gala_amd", "autotune.json")
    return p or None


def _load():
    p = _cache_path()
    if not p or not os.path.exists(p):
        return {}
    try:
        with open(p, encoding="utf-8") as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(key, entry):
    _MEM[key] = entry
    p = _cache_path()
    if not p:
        return
    d = _load()
    d[key] = entry
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + f".{os.getpid()}.tmp"
        with open(tmp, "w", encoding="utf-8") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass  # a read-only home: keep the in-memory result


def key_dop853(potential, N, rtol, atol, dense=False):
    t = dev.torch()
    name = t.cuda.get_device_name(0) if t.cuda.is_available() else "nogpu"
    comps = ",".join(str(c["type"]) for c in potential._components())
    return f"{name}|dop853|pot[{comps}]|N={int(N)}|rtol={rtol:g}|atol={atol:g}|dense={int(dense)}"


def lookup_dop853(potential, N, rtol, atol, dense=False):
    """(block, n_tail) remembered for this problem, or None."""
    k = key_dop853(potential, N, rtol, atol, dense)
    e = _MEM.get(k) or _load().get(k)
    return None if e is None else (int(e["block"]), int(e["n_tail"]))


def _time_ms(fn, repeats):
    """mean of `repeats` timed calls (pass times of the persistent kernel
    scatter with the dynamic queue, and some shapes now and then run 30 %
    slower; the mean rewards a shape that is fast every time)"""
    t = dev.torch()
    s = t.cuda.current_stream()
    ts = []
    for _ in range(repeats):
        e0 = t.cuda.Event(enable_timing=True)
        e1 = t.cuda.Event(enable_timing=True)
        e0.record(s)
        fn()
        e1.record(s)
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    return sum(ts) / len(ts)


def tune_dop853(potential, w0, t0, t1, h0=None, rtol=1e-10, atol=1e-10, order=None,
                blocks=BLOCKS, tails=TAILS, repeats=1, force=False, log=None):
    """Pick (block, n_tail) for gamd_dop853 on this workload by timing whole
    passes.  w0 etc. as for integrate.dop853 (device tensors keep the trials
    free of host copies).  Returns dict(block, n_tail, trials_ms)."""
    from .integrate import dop853
    N = w0.shape[1]
    k = key_dop853(potential, N, rtol, atol)
    if not force:
        e = _MEM.get(k) or _load().get(k)
        if e is not None:
            return dict(e, cached=True)

    def run(b, nt):
        return _time_ms(lambda: dop853(potential, w0, t0, t1, h0, rtol=rtol, atol=atol,
                                       order=order, n_tail=nt, block=b), repeats)

    tail0 = DEFAULT_TAIL if DEFAULT_TAIL in tails else tails[0]
    run(blocks[0], tail0)  # first-launch costs out of the trials
    trials = {}
    for b in blocks:
        trials[f"block={b},n_tail={tail0}"] = ms = run(b, tail0)
        if log:
            log(f"autotune dop853: block {b}, n_tail {tail0}: {ms:.1f} ms")
    best_b = min(blocks, key=lambda b: trials[f"block={b},n_tail={tail0}"])
    for nt in tails:
        key = f"block={best_b},n_tail={nt}"
        if key not in trials:
            trials[key] = ms = run(best_b, nt)
            if log:
                log(f"autotune dop853: block {best_b}, n_tail {nt}: {ms:.1f} ms")
    best_t = min(tails, key=lambda nt: trials[f"block={best_b},n_tail={nt}"])
    entry = {"block": int(best_b), "n_tail": int(best_t), "trials_ms": trials}
    _store(k, entry)
    return dict(entry, cached=False)


# ---------------------------------------------------------------------------
# streamed leapfrog output (gamd_leapfrog_stream): rows per chunk and graph
# capture are speed knobs (identical results); the trials run the caller's
# own integration into its own host buffer

CHUNK_MB = (8, 32, 128, 512)   # candidate chunk sizes (MB of saved rows)


def chunk_rows_for(N, nsave, mb):
    """rows per chunk holding about `mb` MB of saved rows (6 doubles per
    orbit and row), at least 1, at most nsave"""
    rows = int(mb * 1e6 // (48 * max(int(N), 1)))
    return max(1, min(int(nsave), rows))


def default_chunk_rows(N, nsave):
    """untuned choice: ~16 chunks of at least 8 MB (enough chunks for the
    copies to overlap the kernels, few enough launches to stay cheap)"""
    return max(chunk_rows_for(N, nsave, 8), -(-int(nsave) // 16))


def key_stream(potential, N, nsave, stride):
    t = dev.torch()
    name = t.cuda.get_device_name(0) if t.cuda.is_available() else "nogpu"
    comps = ",".join(str(c["type"]) for c in potential._components())
    return f"{name}|leapfrog_stream|pot[{comps}]|N={int(N)}|nsave={int(nsave)}|stride={int(stride)}"


def lookup_stream(potential, N, nsave, stride):
    """(chunk_rows, graph) remembered for this problem, or None."""
    k = key_stream(potential, N, nsave, stride)
    e = _MEM.get(k) or _load().get(k)
    return None if e is None else (int(e["chunk_rows"]), bool(e["graph"]))


def tune_leapfrog_stream(potential, w0, dt, n_steps, save_stride=1, mbs=CHUNK_MB, repeats=2,
                         force=False, log=None):
    """Pick (chunk_rows, graph) for gamd_leapfrog_stream by timing whole runs
    of this integration (wall clock: the call returns when the host buffer is
    complete).  Returns dict(chunk_rows, graph, trials_ms)."""
    import time
    from .integrate import leapfrog_to_host
    N = w0.shape[1]
    nsave = int(n_steps) // int(save_stride) + 1
    k = key_stream(potential, N, nsave, save_stride)
    if not force:
        e = _MEM.get(k) or _load().get(k)
        if e is not None:
            return dict(e, cached=True)
    t = dev.torch()
    buf = t.empty((6, nsave, N), dtype=t.float64, pin_memory=True)
    cands = sorted({chunk_rows_for(N, nsave, mb) for mb in mbs})
    trials = {}
    leapfrog_to_host(potential, w0, dt, min(int(n_steps), int(save_stride)), save_stride,
                     chunk_rows=1)  # first-launch costs out of the trials
    for rows in cands:
        for graph in (False, True):
            ts = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                leapfrog_to_host(potential, w0, dt, n_steps, save_stride, chunk_rows=rows,
                                 graph=graph, out=buf)
                ts.append((time.perf_counter() - t0) * 1e3)
            trials[f"chunk_rows={rows},graph={int(graph)}"] = ms = sum(ts) / len(ts)
            if log:
                log(f"autotune leapfrog stream: {rows} rows per chunk, graph {int(graph)}: "
                    f"{ms:.1f} ms")
    best = min(trials, key=trials.get)
    rows = int(best.split(",")[0].split("=")[1])
    graph = best.endswith("graph=1")
    entry = {"chunk_rows": rows, "graph": graph, "trials_ms": trials}
    _store(k, entry)
    return dict(entry, cached=False)


# ---------------------------------------------------------------------------
# streamed DOP853 dense output (gamd_dop853_dense_stream): samples per window
# and graph capture are speed knobs (identical results)


def key_dense(potential, N, n_out):
    t = dev.torch()
    name = t.cuda.get_device_name(0) if t.cuda.is_available() else "nogpu"
    comps = ",".join(str(c["type"]) for c in potential._components())
    return f"{name}|dop853_dense|pot[{comps}]|N={int(N)}|n_out={int(n_out)}"


def lookup_dense(potential, N, n_out):
    """(window_rows, graph) remembered for this problem, or None."""
    k = key_dense(potential, N, n_out)
    e = _MEM.get(k) or _load().get(k)
    return None if e is None else (int(e["window_rows"]), bool(e["graph"]))


def tune_dop853_dense(potential, w0, t0, t1, t_out, h0=None, rtol=1e-10, atol=1e-10,
                      order=None, n_tail=0, mbs=CHUNK_MB, repeats=2, force=False, log=None):
    """Pick (window_rows, graph) for gamd_dop853_dense_stream by timing whole
    runs of this integration (wall clock: the call returns when the host
    buffer is complete).  Returns dict(window_rows, graph, trials_ms)."""
    import time
    from .integrate import dop853_to_host
    N = w0.shape[1]
    n_out = len(t_out)
    k = key_dense(potential, N, n_out)
    if not force:
        e = _MEM.get(k) or _load().get(k)
        if e is not None:
            return dict(e, cached=True)
    t = dev.torch()
    buf = t.empty((6, n_out, N), dtype=t.float64, pin_memory=True)
    cands = {chunk_rows_for(N, n_out, mb) for mb in mbs}
    # the whole grid as one window (every sample in HBM, one copy at the end)
    # when its workspace takes at most half the free device memory: windows
    # restart the persistent launch, and where the longest orbits set each
    # window's time one launch can beat the overlapped copies
    from . import _lib
    whole = int(_lib.load().gamd_dop853_dense_workspace_bytes(N, n_out))
    if whole <= t.cuda.mem_get_info()[0] // 2:
        cands.add(n_out)
    cands = sorted(cands)
    trials = {}
    for rows in cands:
        for graph in (False, True):
            ts = []
            for _ in range(repeats):
                t_start = time.perf_counter()
                dop853_to_host(potential, w0, t0, t1, t_out, h0=h0, rtol=rtol, atol=atol,
                               order=order, n_tail=n_tail, window_rows=rows, graph=graph,
                               out=buf)
                ts.append((time.perf_counter() - t_start) * 1e3)
            trials[f"window_rows={rows},graph={int(graph)}"] = ms = sum(ts) / len(ts)
            if log:
                log(f"autotune dop853 dense output: {rows} samples per window, graph "
                    f"{int(graph)}: {ms:.1f} ms")
    best = min(trials, key=trials.get)
    rows = int(best.split(",")[0].split("=")[1])
    graph = best.endswith("graph=1")
    entry = {"window_rows": rows, "graph": graph, "trials_ms": trials}
    _store(k, entry)
    return dict(entry, cached=False)
