"""Launch-shape autotuner for the persistent DOP853 kernel.

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
