"""Time grids, integrator lookup and the orbit-integration dispatch.

Mirrors gala's src/gala/integrate/timespec.py:10-147 (parse_time_specification),
integrate/lookup.py:14-84 (get_integrator) and the Cython-path dispatch of
Hamiltonian.integrate_orbit (potential/hamiltonian/chamiltonian.pyx:274-380).
The integrations themselves run in the HIP kernels (gamd_leapfrog,
gamd_dop853).  Lower-level functions taking device tensors (``leapfrog``,
``dop853``) serve callers whose inputs already live in HBM.
"""
import numpy as np

from . import _device as dev
from . import _lib

__all__ = ["parse_time_specification", "get_integrator", "LeapfrogIntegrator",
           "DOPRI853Integrator", "Ruth4Integrator", "RK5Integrator", "leapfrog", "ruth4", "dop853",
           "dop853_batched", "leapfrog_to_host", "dop853_to_host", "pilot_order",
           "nbody_leapfrog", "nbody_ruth4", "nbody_dop853", "integrate_hamiltonian"]


class Integrator:
    """Integrator name tags (gala's classes live in integrate/pyintegrators;
    here they only select the HIP kernel)."""
    name = None


class LeapfrogIntegrator(Integrator):
    name = "leapfrog"


class DOPRI853Integrator(Integrator):
    name = "dopri853"


class Ruth4Integrator(Integrator):
    name = "ruth4"


class RK5Integrator(Integrator):
    name = "rk5"


_integrator_name_mapping = {
    "leapfrog": LeapfrogIntegrator,
    "dopri853": DOPRI853Integrator,
    "dop853": DOPRI853Integrator,
    "rk5": RK5Integrator,
    "ruth4": Ruth4Integrator,
}


def get_integrator(name):
    """Integrator class from a (case-insensitive) name or class (lookup.py:23-84)."""
    if not isinstance(name, str):
        if name not in list(_integrator_name_mapping.values()):
            raise ValueError(f"Integrator class '{name}' is not recognized. Valid classes are: "
                             f"{list(_integrator_name_mapping.values())}")
        return name
    try:
        return _integrator_name_mapping[name.lower()]
    except KeyError as e:
        raise ValueError(f"Integrator name '{name}' is not recognized. Valid names are: "
                         f"{list(_integrator_name_mapping.keys())}") from e


def parse_time_specification(units=None, dt=None, n_steps=None, t1=None, t2=None, t=None):
    """Array of times from one of: (dt, n_steps[, t1]), (dt, t1, t2),
    (dt array, t1), (n_steps, t1, t2) or t (timespec.py:10-147)."""
    if n_steps is not None:
        n_steps = int(n_steps)
    if t is not None:
        return np.asarray(t).astype(np.float64)
    if dt is None and (t1 is None or t2 is None or n_steps is None):
        raise ValueError("Invalid specification of integration time. See docstring for more "
                         "information.")
    if dt is not None and n_steps is not None:
        if t1 is None:
            t1 = 0.0
        return parse_time_specification(units, dt=np.ones(n_steps + 1) * dt, t1=t1)
    if dt is not None and t1 is not None and t2 is not None:
        if t2 < t1 and dt < 0:
            t_i, times = t1, []
            while t_i > t2:
                times.append(t_i)
                t_i += dt
            if times[-1] != t2:
                times.append(t2)
            return np.array(times, dtype=np.float64)
        if t2 > t1 and dt > 0:
            t_i, times = t1, []
            while t_i < t2:
                times.append(t_i)
                t_i += dt
            return np.array(times, dtype=np.float64)
        if dt == 0:
            raise ValueError("dt must be non-zero.")
        raise ValueError("If t2 < t1, dt must be negative. If t1 < t2, dt must be positive.")
    if isinstance(dt, np.ndarray) and t1 is not None:
        times = np.cumsum(np.append([0.0], dt)) + t1
        return times[:-1].astype(np.float64)
    if dt is None and not (t1 is None or t2 is None or n_steps is None):
        return np.linspace(t1, t2, n_steps, endpoint=True).astype(np.float64)
    raise ValueError("Invalid options. See docstring.")


# ---------------------------------------------------------------------------
# device-level calls (inputs and outputs are device tensors)

def leapfrog(potential, w0, dt, n_steps, save_stride=1, save=True, n_steps_per_orbit=None,
             t0=0.0):
    """Leapfrog on device.  w0: (6, N) float64 device tensor.
    Returns (w_out (6, n_steps//save_stride + 1, N) or None, w_final (6, N))."""
    lib = _lib.load()
    w0 = dev.phase_space(w0)
    N = w0.shape[1]
    wout = dev.empty((6, int(n_steps) // int(save_stride) + 1, N)) if save else None
    wfin = dev.empty((6, N))
    nsp = None
    if n_steps_per_orbit is not None:
        nsp = dev.per_orbit(n_steps_per_orbit, N, "n_steps_per_orbit", dtype=dev.torch().int64)
        n_steps = int(nsp.max().item()) if N else 0
    _lib.check(lib.gamd_leapfrog(potential.c_handle, w0.data_ptr(), N, float(t0), float(dt),
                                 int(n_steps), dev.ptr(nsp), int(save_stride), dev.ptr(wout),
                                 wfin.data_ptr(), dev.stream_ptr()), "gamd_leapfrog")
    return wout, wfin


def leapfrog_to_host(potential, w0, dt, n_steps, save_stride=1, t0=0.0, chunk_rows=None,
                     graph=None, out=None):
    """Leapfrog with the saved trajectory streamed to pinned host memory while
    the integration runs (gamd_leapfrog_stream): chunks of chunk_rows saved
    rows are computed and copied on two streams, so the copies overlap the
    kernels.  chunk_rows / graph: None = the autotuner's remembered choice
    for this problem, else a default of ~16 chunks.  out: a pinned host
    float64 tensor (6, nsave, N) to reuse.  Returns (w_out numpy view of the
    pinned buffer (6, nsave, N), w_final (6, N) device), equal bit for bit to
    leapfrog(..., save_stride)."""
    from . import autotune
    t = dev.require_gpu()
    lib = _lib.load()
    w0 = dev.phase_space(w0)
    N = w0.shape[1]
    n_steps, save_stride = int(n_steps), int(save_stride)
    if n_steps < 0 or save_stride < 1:
        raise ValueError("n_steps must be >= 0 and save_stride >= 1")
    nsave = n_steps // save_stride + 1
    if chunk_rows is None or graph is None:
        tuned = autotune.lookup_stream(potential, N, nsave, save_stride)
        if chunk_rows is None:
            chunk_rows = tuned[0] if tuned else autotune.default_chunk_rows(N, nsave)
        if graph is None:
            graph = tuned[1] if tuned else False
    chunk_rows = max(1, min(int(chunk_rows), nsave))
    if out is None or tuple(out.shape) != (6, nsave, N):
        out = t.empty((6, nsave, N), dtype=t.float64, pin_memory=True)
    wfin = dev.empty((6, N))
    nbytes = int(lib.gamd_leapfrog_stream_workspace_bytes(N, chunk_rows))
    ws = t.empty(max(nbytes, 8), dtype=t.uint8, device="cuda")
    _lib.check(lib.gamd_leapfrog_stream(potential.c_handle, w0.data_ptr(), N, float(t0),
                                        float(dt), n_steps, save_stride, chunk_rows,
                                        int(bool(graph)), out.data_ptr(), wfin.data_ptr(),
                                        ws.data_ptr(), nbytes, dev.stream_ptr()),
               "gamd_leapfrog_stream")
    return out.numpy(), wfin


def ruth4(potential, w0, dt, n_steps, save_stride=1, save=True, t0=0.0):
    """Ruth4 on device (ruth4.pyx:24-106).  w0: (6, N) float64.
    Returns (w_out (6, n_steps//save_stride + 1, N) or None, w_final (6, N))."""
    lib = _lib.load()
    w0 = dev.phase_space(w0)
    N = w0.shape[1]
    wout = dev.empty((6, int(n_steps) // int(save_stride) + 1, N)) if save else None
    wfin = dev.empty((6, N))
    _lib.check(lib.gamd_ruth4(potential.c_handle, w0.data_ptr(), N, float(t0), float(dt),
                              int(n_steps), int(save_stride), dev.ptr(wout), wfin.data_ptr(),
                              dev.stream_ptr()), "gamd_ruth4")
    return wout, wfin


# 1 % of each orbit's span ranks the 1M-orbit uneven workload as well as 2 % (full
# launch 114.3-114.6 ms either way) in half the pilot time (3.2 against 5.7-6.1 ms);
# at 0.5 % the ranking starts to cost the full launch
PILOT_FRACTION = 0.01
# integrate_orbit orders per-orbit DOP853 runs of at least this many orbits by a pilot pass
PILOT_MIN_ORBITS = 4096


def pilot_order(potential, w0, t0, t1, h0=None, rtol=1e-10, atol=1e-10, nmax=0, nstiff=0,
                hmax=0.0, fraction=PILOT_FRACTION, block=0):
    """Longest-first queue order for gamd_dop853 from a pilot pass: every
    orbit integrated over the first `fraction` of its time span with the same
    tolerances; its step count / fraction estimates the whole run's, and the
    orbits are ordered by that estimate, descending (device int32 tensor).
    Eccentric orbits through the nucleus cost many more steps per orbital
    period than their dynamical time suggests; the pilot measures it."""
    t = dev.require_gpu()
    w0 = dev.phase_space(w0)
    N = w0.shape[1]
    t0d, t1d = dev.per_orbit(t0, N, "t0"), dev.per_orbit(t1, N, "t1")
    rp = dop853(potential, w0, t0d, t0d + (t1d - t0d) * fraction, h0, rtol=rtol, atol=atol,
                nmax=nmax, nstiff=nstiff, hmax=hmax, block=block)
    est = rp["stats"][:, 0].to(t.float64)
    return t.argsort(est, descending=True, stable=True).to(t.int32)


def dop853(potential, w0, t0, t1, h0=None, rtol=1e-10, atol=1e-10, nmax=0, nstiff=0, hmax=0.0,
           t_out=None, order=None, n_tail=0, block=0):
    """Per-orbit adaptive DOP853 on device.  w0 (6, N); t0, t1, h0 per orbit
    (scalars broadcast); order: queue order (a permutation, or "pilot" for
    pilot_order's estimate, computed in the same call); n_tail: the first
    n_tail orbits of that order run on dedicated 16-lane rows from the start.
    Returns dict(w_final, status, stats[, w_out])."""
    N = dev.check_w0(w0)
    for v, name in ((t0, "t0"), (t1, "t1"), (h0, "h0")):
        dev.check_per_orbit(v, N, name)
    if order is not None and not isinstance(order, str):
        dev.check_permutation(order, N)
    t = dev.require_gpu()
    lib = _lib.load()
    w0 = dev.phase_space(w0)
    if isinstance(order, str):
        if order != "pilot":
            raise ValueError(f"unknown order {order!r}")
        order = pilot_order(potential, w0, t0, t1, h0, rtol=rtol, atol=atol, nmax=nmax,
                            nstiff=nstiff, hmax=hmax, block=block)
    t0d, t1d = dev.per_orbit(t0, N, "t0"), dev.per_orbit(t1, N, "t1")
    h0d = dev.per_orbit(h0, N, "h0")
    wfin = dev.empty((6, N))
    status = dev.empty((N,), dtype=t.int32)
    stats = dev.empty((N, 4), dtype=t.int64)
    wout = tout = None
    if t_out is not None:
        tout = dev.to_dev(t_out)
        wout = dev.empty((6, tout.shape[0], N))
    ordd = dev.to_dev(order, dtype=t.int32) if order is not None else None  # checked above
    _lib.check(lib.gamd_dop853(potential.c_handle, w0.data_ptr(), N, t0d.data_ptr(),
                               t1d.data_ptr(), dev.ptr(h0d), float(rtol), float(atol), int(nmax),
                               int(nstiff), float(hmax), dev.ptr(tout),
                               0 if tout is None else tout.shape[0], dev.ptr(wout),
                               wfin.data_ptr(), status.data_ptr(), stats.data_ptr(),
                               dev.ptr(ordd), int(n_tail), int(block), dev.stream_ptr()),
               "gamd_dop853")
    out = {"w_final": wfin, "status": status, "stats": stats}
    if wout is not None:
        out["w_out"] = wout
    return out


def dop853_to_host(potential, w0, t0, t1, t_out, h0=None, rtol=1e-10, atol=1e-10, nmax=0,
                   nstiff=0, hmax=0.0, order=None, n_tail=0, block=0, window_rows=None,
                   graph=None, out=None):
    """dop853(..., t_out=t_out) with the dense output streamed to pinned host
    memory while the integration runs (gamd_dop853_dense_stream): the t_out
    grid is taken in windows of window_rows samples, each window's samples
    copied on a second stream while the next window integrates.  window_rows
    / graph: None = the autotuner's remembered choice for this problem, else
    a default of ~16 windows.  order, n_tail, block: as dop853 (speed knobs).
    out: a pinned host float64 tensor (6, n_out, N) to reuse.  Returns
    dict(w_out = numpy view of the pinned buffer (6, n_out, N), w_final,
    status, stats (device)), equal bit for bit to dop853's."""
    from . import autotune
    N = dev.check_w0(w0)
    for v, name in ((t0, "t0"), (t1, "t1"), (h0, "h0")):
        dev.check_per_orbit(v, N, name)
    if order is not None and not isinstance(order, str):
        dev.check_permutation(order, N)
    t = dev.require_gpu()
    lib = _lib.load()
    w0 = dev.phase_space(w0)
    if isinstance(order, str):
        if order != "pilot":
            raise ValueError(f"unknown order {order!r}")
        order = pilot_order(potential, w0, t0, t1, h0, rtol=rtol, atol=atol, nmax=nmax,
                            nstiff=nstiff, hmax=hmax, block=block)
    tout = dev.to_dev(t_out)
    if tout.ndim != 1 or tout.shape[0] < 1:
        raise ValueError("t_out must be a non-empty 1-d array of times")
    n_out = int(tout.shape[0])
    if window_rows is None or graph is None:
        tuned = autotune.lookup_dense(potential, N, n_out)
        if window_rows is None:
            window_rows = tuned[0] if tuned else autotune.default_chunk_rows(N, n_out)
        if graph is None:
            graph = tuned[1] if tuned else False
    window_rows = max(1, min(int(window_rows), n_out))
    t0d, t1d = dev.per_orbit(t0, N, "t0"), dev.per_orbit(t1, N, "t1")
    h0d = dev.per_orbit(h0, N, "h0")
    if out is None or tuple(out.shape) != (6, n_out, N):
        out = t.empty((6, n_out, N), dtype=t.float64, pin_memory=True)
    wfin = dev.empty((6, N))
    status = dev.empty((N,), dtype=t.int32)
    stats = dev.empty((N, 4), dtype=t.int64)
    ordd = dev.to_dev(order, dtype=t.int32) if order is not None else None  # checked above
    nbytes = int(lib.gamd_dop853_dense_workspace_bytes(N, window_rows))
    ws = t.empty(max(nbytes, 8), dtype=t.uint8, device="cuda")
    _lib.check(lib.gamd_dop853_dense_stream(potential.c_handle, w0.data_ptr(), N,
                                            t0d.data_ptr(), t1d.data_ptr(), dev.ptr(h0d),
                                            float(rtol), float(atol), int(nmax), int(nstiff),
                                            float(hmax), tout.data_ptr(), n_out, window_rows,
                                            int(bool(graph)), out.data_ptr(), wfin.data_ptr(),
                                            status.data_ptr(), stats.data_ptr(), dev.ptr(ordd),
                                            int(n_tail), int(block), ws.data_ptr(), nbytes,
                                            dev.stream_ptr()),
               "gamd_dop853_dense_stream")
    return {"w_out": out.numpy(), "w_final": wfin, "status": status, "stats": stats}


MAX_NBATCH = 1024  # GAMD_DOP853_MAX_NBATCH: orbits of one batched system


def dop853_batched(potential, w0, t0, t1, h0, nbatch=100, rtol=1e-10, atol=1e-10, nmax=0,
                   nstiff=0, hmax=0.0, t_out=None):
    """DOP853 with gala's default batching (dop853_integrate_hamiltonian,
    dop853.pyx:196-250): orbits [i, i + nbatch) of w0 (6, N) form one
    (6, nb) system with one shared adaptive step, integrated from t0 to t1
    (scalars) with initial step h0 (gala passes t[1] - t[0]); dense output at
    t_out.  1 <= nbatch <= MAX_NBATCH.  Returns dict(w_final, status, stats
    [, w_out]) with each orbit's batch's status and step counts."""
    N = dev.check_w0(w0)
    nbatch = int(nbatch)
    if nbatch < 1:
        raise ValueError(f"nbatch must be at least 1, got {nbatch}")
    if nbatch > MAX_NBATCH:
        raise NotImplementedError(f"nbatch={nbatch}: one batched system holds at most "
                                  f"{MAX_NBATCH} orbits on the GPU (one workgroup)")
    t = dev.require_gpu()
    lib = _lib.load()
    w0 = dev.phase_space(w0)
    wfin = dev.empty((6, N))
    status = dev.empty((N,), dtype=t.int32)
    stats = dev.empty((N, 4), dtype=t.int64)
    wout = tout = None
    if t_out is not None:
        tout = dev.to_dev(t_out)
        wout = dev.empty((6, tout.shape[0], N))
    _lib.check(lib.gamd_dop853_batched(potential.c_handle, w0.data_ptr(), N, float(t0),
                                       float(t1), float(h0), float(rtol), float(atol), int(nmax),
                                       int(nstiff), float(hmax), nbatch, dev.ptr(tout),
                                       0 if tout is None else tout.shape[0], dev.ptr(wout),
                                       wfin.data_ptr(), status.data_ptr(), stats.data_ptr(),
                                       dev.stream_ptr()), "gamd_dop853_batched")
    out = {"w_final": wfin, "status": status, "stats": stats}
    if wout is not None:
        out["w_out"] = wout
    return out


def _bodies(body_potentials):
    """Host array of body handles (None = massless) and the keepalive list."""
    handles = [None if p is None else p.c_handle for p in body_potentials]
    return _lib.body_array(handles), len(handles)


def nbody_leapfrog(potential, body_potentials, w0, dt, n_steps, save_stride=1, save=True,
                   t0=0.0):
    """leapfrog_integrate_nbody on device (leapfrog.pyx:161-257).
    body_potentials: potentials of bodies 0..K-1 (None = massless), each
    centred on its body; bodies K.. of w0 (6, n) are test particles.
    Returns (w_out (6, n_steps//save_stride + 1, n) or None, w_final (6, n))."""
    lib = _lib.load()
    w0 = dev.phase_space(w0)
    n = w0.shape[1]
    arr, K = _bodies(body_potentials)
    wout = dev.empty((6, int(n_steps) // int(save_stride) + 1, n)) if save else None
    wfin = dev.empty((6, n))
    _lib.check(lib.gamd_nbody_leapfrog(potential.c_handle, arr, K, w0.data_ptr(), n, float(t0),
                                       float(dt), int(n_steps), int(save_stride), dev.ptr(wout),
                                       wfin.data_ptr(), dev.stream_ptr()), "gamd_nbody_leapfrog")
    return wout, wfin


def nbody_ruth4(potential, body_potentials, w0, dt, n_steps, save_stride=1, save=True, t0=0.0):
    """ruth4_integrate_nbody on device (ruth4.pyx:138-261); arguments and
    results as nbody_leapfrog."""
    lib = _lib.load()
    w0 = dev.phase_space(w0)
    n = w0.shape[1]
    arr, K = _bodies(body_potentials)
    wout = dev.empty((6, int(n_steps) // int(save_stride) + 1, n)) if save else None
    wfin = dev.empty((6, n))
    _lib.check(lib.gamd_nbody_ruth4(potential.c_handle, arr, K, w0.data_ptr(), n, float(t0),
                                    float(dt), int(n_steps), int(save_stride), dev.ptr(wout),
                                    wfin.data_ptr(), dev.stream_ptr()), "gamd_nbody_ruth4")
    return wout, wfin


def nbody_dop853(potential, body_potentials, w0, t0, t1, h0, rtol=1e-10, atol=1e-10, nmax=0,
                 nstiff=-1, hmax=0.0, t_out=None):
    """direct_nbody_dop853 on device (nbody.pyx:30-115): the n bodies of w0
    (6, n) as one system with one shared adaptive step.
    Returns dict(w_final (6, n), status (1,), stats (1, 4)[, w_out (6, n_out, n)])."""
    t = dev.require_gpu()
    lib = _lib.load()
    w0 = dev.phase_space(w0)
    n = w0.shape[1]
    arr, K = _bodies(body_potentials)
    wfin = dev.empty((6, n))
    status = dev.empty((1,), dtype=t.int32)
    stats = dev.empty((1, 4), dtype=t.int64)
    wout = tout = None
    if t_out is not None:
        tout = dev.to_dev(t_out)
        wout = dev.empty((6, tout.shape[0], n))
    _lib.check(lib.gamd_nbody_dop853(potential.c_handle, arr, K, w0.data_ptr(), n, float(t0),
                                     float(t1), float(h0), float(rtol), float(atol), int(nmax),
                                     int(nstiff), float(hmax), dev.ptr(tout),
                                     0 if tout is None else tout.shape[0], dev.ptr(wout),
                                     wfin.data_ptr(), status.data_ptr(), stats.data_ptr(),
                                     dev.stream_ptr()), "gamd_nbody_dop853")
    out = {"w_final": wfin, "status": status, "stats": stats}
    if wout is not None:
        out["w_out"] = wout
    return out


# ---------------------------------------------------------------------------
# Hamiltonian.integrate_orbit

def _prepare_w0(hamiltonian, w0):
    from .dynamics import PhaseSpacePosition
    if isinstance(w0, PhaseSpacePosition):
        arr = w0.w(hamiltonian.units)
    else:
        arr = np.asarray(w0, dtype=np.float64)
    if arr.shape[0] != hamiltonian.ndim:
        raise ValueError(f"Invalid initial conditions shape {arr.shape}. Expected shape "
                         f"({hamiltonian.ndim}, ...) but got shape {arr.shape}.")
    orig_shape = arr.shape
    return orig_shape, np.ascontiguousarray(arr.reshape(arr.shape[0], -1))


def integrate_hamiltonian(hamiltonian, w0, Integrator=None, Integrator_kwargs=None,
                          save_all=True, **time_spec):
    """The dispatch of chamiltonian.pyx:274-380 onto the HIP kernels.

    Leapfrog: dt = t[1]-t[0], len(t)-1 steps (leapfrog.pyx:78, 106).
    DOP853: orbits are integrated from t[0] to t[-1] with h0 = t[1]-t[0]
    (dop853.pyx:171) in batches of nbatch (default 100, chamiltonian.pyx:348)
    orbits sharing one adaptive step (dop853.pyx:228-245); nbatch=1 makes
    every orbit its own system.  With save_all the dense output samples every
    t (dop853.cpp:584-612).  Integrator_kwargs: atol, rtol (default 1e-10),
    nmax, err_if_fail (default 1, raise on a negative status), nbatch, and,
    beyond gala's keys, nstiff and dt_max.  log_output is accepted and
    ignored (the kernels print nothing).
    """
    from .dynamics import Orbit
    Integrator_kwargs = dict(Integrator_kwargs or {})
    if Integrator is None:
        Integrator = LeapfrogIntegrator
    Integrator = get_integrator(Integrator)
    orig_shape, arr = _prepare_w0(hamiltonian, w0)
    t = np.ascontiguousarray(parse_time_specification(hamiltonian.units, **time_spec))
    pot = hamiltonian.potential
    if Integrator is LeapfrogIntegrator:
        if len(t) < 2:
            raise ValueError("need at least two times")
        dt = t[1] - t[0]
        if save_all:
            # the trajectory streams to pinned host memory while it is computed
            w, _ = leapfrog_to_host(pot, arr, dt, len(t) - 1, save_stride=1, t0=t[0])
        else:
            _, wfin = leapfrog(pot, arr, dt, len(t) - 1, save=False, t0=t[0])
            w = dev.to_host(wfin)
    elif Integrator is Ruth4Integrator:
        if len(t) < 2:
            raise ValueError("need at least two times")
        wout, wfin = ruth4(pot, arr, t[1] - t[0], len(t) - 1, save_stride=1,
                           save=bool(save_all), t0=t[0])
        w = dev.to_host(wout) if save_all else dev.to_host(wfin)
    elif Integrator is DOPRI853Integrator:
        if len(t) < 2:
            raise ValueError("need at least two times")
        kw = dict(rtol=Integrator_kwargs.get("rtol", 1e-10),
                  atol=Integrator_kwargs.get("atol", 1e-10),
                  nmax=Integrator_kwargs.get("nmax", 0),
                  nstiff=Integrator_kwargs.get("nstiff", 0),
                  hmax=Integrator_kwargs.get("dt_max", 0.0),
                  t_out=t if save_all else None)
        nbatch = int(Integrator_kwargs.get("nbatch", 100))  # chamiltonian.pyx:348
        if nbatch == 1:
            # many orbits: longest first from a pilot pass, the longest on
            # 16-lane rows (the same results for any order and n_tail)
            N = arr.shape[1]
            if N >= PILOT_MIN_ORBITS:
                kw.update(order="pilot", n_tail=min(1024, N // 64))
        if nbatch == 1 and save_all:
            # the dense output streams to pinned host memory while it is computed
            kw.pop("t_out")
            res = dop853_to_host(pot, arr, t[0], t[-1], t, h0=t[1] - t[0], **kw)
        elif nbatch == 1:
            res = dop853(pot, arr, t[0], t[-1], t[1] - t[0], **kw)
        else:
            res = dop853_batched(pot, arr, t[0], t[-1], t[1] - t[0], nbatch=nbatch, **kw)
        status = dev.to_host(res["status"])
        if int(Integrator_kwargs.get("err_if_fail", 1)) == 1 and np.any(status < 0):
            bad = int(status[status < 0][0])
            raise RuntimeError(f"Integration failed with code {bad}")
        if save_all:
            w = res["w_out"] if nbatch == 1 else dev.to_host(res["w_out"])
        else:
            w = dev.to_host(res["w_final"])
    else:
        raise ValueError(f"Cython integration not supported for '{Integrator!r}'")
    if w.shape[-1] == 1:  # chamiltonian.pyx:355-356
        w = w[..., 0]
    if not save_all:
        w = w[:, None]
        t = t[-1:]
    return Orbit.from_w(w=w, units=hamiltonian.units, t=t, hamiltonian=hamiltonian)
