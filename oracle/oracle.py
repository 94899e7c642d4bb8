"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker / the CPU baseline.  The product
(gala_amd/) never imports it.

The C side (gala_oracle.c) restates gala's hot-path algorithms; this file
only marshals numpy arrays.  Potentials are described independently of the
product's classes: a list of components, each built by one of the helpers
below from raw parameter values, so that a packing mistake in the product
cannot be mirrored here.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAXC = 16
NULL, HERNQUIST, PLUMMER, MIYAMOTONAGAI, NFW, SCF = range(6)

# G in kpc^3 / (Msun Myr^2): CODATA-2018 G times IAU-2015 nominal solar mass,
# the value astropy's constants give in gala's galactic unit system.
G_GALACTIC = 4.498502151469553e-12


class Comp(C.Structure):
    _fields_ = [("type", C.c_int), ("shift_rotate", C.c_int), ("nmax", C.c_int),
                ("lmax", C.c_int), ("pars", C.c_double * 8), ("q0", C.c_double * 3),
                ("R", C.c_double * 9), ("S", C.c_void_p), ("T", C.c_void_p)]


class Pot(C.Structure):
    _fields_ = [("n", C.c_int), ("pad", C.c_int), ("c", Comp * MAXC)]


def _lib(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library {path} is not built (run make -C oracle)")
    return C.CDLL(path)


_LIBS = {}


def lib(fast=False):
    key = "liboracle_fast.so" if fast else "liboracle.so"
    if key not in _LIBS:
        L = _lib(key)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        lp = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        P = C.POINTER(Pot)
        L.orc_value.restype = C.c_double
        L.orc_value.argtypes = [P, C.c_double, dp]
        L.orc_density.restype = C.c_double
        L.orc_density.argtypes = [P, C.c_double, dp]
        L.orc_hessian.argtypes = [P, C.c_double, dp, dp]
        L.orc_gradient.argtypes = [P, C.c_size_t, C.c_double, dp, dp]
        L.orc_d2_dr2.restype = C.c_double
        L.orc_d2_dr2.argtypes = [P, C.c_double, dp]
        L.orc_leapfrog.argtypes = [P, dp, C.c_size_t, dp, C.c_size_t, C.c_int, dp]
        L.orc_ruth4.argtypes = [P, dp, C.c_size_t, dp, C.c_size_t, C.c_int, dp]
        L.orc_dop853.argtypes = [P, dp, C.c_size_t, dp, C.c_size_t, C.c_double, C.c_double,
                                 C.c_long, C.c_long, C.c_double, C.c_int, C.c_int, dp, ip, lp]
        L.orc_dop853_orbits.argtypes = [P, dp, C.c_size_t, dp, dp, dp, C.c_double, C.c_double,
                                        C.c_long, C.c_long, dp, ip, lp]
        L.orc_mockstream_leapfrog.argtypes = [P, C.c_void_p, C.c_int, dp, C.c_size_t, dp,
                                              C.c_size_t, dp, ip, dp, C.c_double, dp, dp]
        L.orc_mockstream_dop853.argtypes = [P, C.c_void_p, C.c_int, dp, C.c_size_t, dp, ip, dp,
                                            C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_long, dp, dp, ip, lp]
        L.orc_nbody_leapfrog.argtypes = [P, C.c_void_p, C.c_int, dp, C.c_int, dp, C.c_size_t,
                                         C.c_int, dp]
        L.orc_nbody_ruth4.argtypes = [P, C.c_void_p, C.c_int, dp, C.c_int, dp, C.c_size_t,
                                      C.c_int, dp]
        L.orc_nbody_dop853.restype = C.c_int
        L.orc_nbody_dop853.argtypes = [P, C.c_void_p, C.c_int, dp, C.c_int, dp, C.c_size_t,
                                       C.c_double, C.c_double, C.c_long, C.c_double, C.c_int, dp,
                                       lp]
        L.orc_fardal_sample.restype = C.c_long
        L.orc_fardal_sample.argtypes = [P, C.c_double, dp, dp, dp, dp, ip, C.c_size_t, C.c_int,
                                        C.c_int, C.c_int, dp, dp, dp, dp]
        L.orc_num_threads.restype = C.c_int
        L.orc_sizeof_pot.restype = C.c_size_t
        assert L.orc_sizeof_pot() == C.sizeof(Pot)
        _LIBS[key] = L
    return _LIBS[key]


def ref_lib():
    """gala's own compiled C++ behind oracle/ref_driver.cpp, or None when the
    reference tree was not available to build it."""
    path = os.path.join(HERE, "_ref", "libgala_ref.so")
    if not os.path.exists(path):
        return None
    if "ref" not in _LIBS:
        L = C.CDLL(path)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        P = C.POINTER(Pot)
        L.ref_value.restype = C.c_double
        L.ref_value.argtypes = [P, C.c_double, dp]
        L.ref_density.restype = C.c_double
        L.ref_density.argtypes = [P, C.c_double, dp]
        L.ref_gradient.argtypes = [P, C.c_size_t, C.c_double, dp, dp]
        L.ref_hessian.argtypes = [P, C.c_double, dp, dp]
        L.ref_d2_dr2.restype = C.c_double
        L.ref_d2_dr2.argtypes = [P, C.c_double, dp]
        L.ref_mockstream_dop853.argtypes = [P, C.c_void_p, C.c_int, dp, C.c_size_t, dp, ip, dp,
                                            C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_long, dp, dp, ip]
        L.ref_nbody_dop853.restype = C.c_int
        L.ref_nbody_dop853.argtypes = [P, C.c_void_p, C.c_int, dp, C.c_int, dp, C.c_size_t,
                                       C.c_double, C.c_double, C.c_long, C.c_double, C.c_int, dp]
        L.ref_nbody_leapfrog.argtypes = [P, C.c_void_p, C.c_int, dp, C.c_int, dp, C.c_size_t,
                                         C.c_int, dp]
        L.ref_nbody_ruth4.argtypes = [P, C.c_void_p, C.c_int, dp, C.c_int, dp, C.c_size_t,
                                      C.c_int, dp]
        L.ref_dop853.argtypes = [P, dp, C.c_size_t, dp, C.c_size_t, C.c_double, C.c_double,
                                 C.c_long, C.c_long, C.c_double, C.c_int, C.c_int, dp, ip]
        _LIBS["ref"] = L
    return _LIBS["ref"]


# --------------------------------------------------------------------------
# component helpers (parameters in gala's declaration order, G first)

def comp(kind, pars, q0=None, R=None):
    c = {"type": kind, "pars": [float(p) for p in pars]}
    q0 = np.zeros(3) if q0 is None else np.asarray(q0, float)
    R = np.eye(3) if R is None else np.asarray(R, float)
    c["q0"] = q0
    c["R"] = R
    c["shift_rotate"] = int(not (np.all(q0 == 0) and np.all(R == np.eye(3))))
    return c


def hernquist(m, c, G=G_GALACTIC, **kw):
    return comp(HERNQUIST, [G, m, c], **kw)


def plummer(m, b, G=G_GALACTIC, **kw):
    return comp(PLUMMER, [G, m, b], **kw)


def miyamotonagai(m, a, b, G=G_GALACTIC, **kw):
    return comp(MIYAMOTONAGAI, [G, m, a, b], **kw)


def nfw(m, r_s, G=G_GALACTIC, **kw):
    return comp(NFW, [G, m, r_s, 1.0, 1.0, 1.0], **kw)


def scf(m, r_s, S, T, G=G_GALACTIC, **kw):
    S = np.ascontiguousarray(S, dtype=np.float64)
    T = np.ascontiguousarray(T, dtype=np.float64)
    c = comp(SCF, [G, m, r_s], **kw)
    c["S"], c["T"] = S, T
    c["nmax"], c["lmax"] = S.shape[0] - 1, S.shape[1] - 1
    return c


def milky_way_like(G=G_GALACTIC):
    """The four-component composite of the benchmark workloads: Hernquist
    nucleus, Hernquist bulge, Miyamoto-Nagai disk, NFW halo (gala's
    MilkyWayPotential v1 layout, builtin/special.py:98-124)."""
    return [hernquist(1.7e9, 0.07, G=G), hernquist(5e9, 1.0, G=G),
            miyamotonagai(6.8e10, 3.0, 0.28, G=G), nfw(5.4e11, 15.6, G=G)]


def make_pot(components):
    P = Pot()
    P.n = len(components)
    keep = []
    for i, c in enumerate(components):
        cc = P.c[i]
        cc.type = c["type"]
        cc.shift_rotate = c["shift_rotate"]
        for k, v in enumerate(c["pars"]):
            cc.pars[k] = v
        for k in range(3):
            cc.q0[k] = c["q0"][k]
        for k, v in enumerate(np.asarray(c["R"], float).ravel()):
            cc.R[k] = v
        if c["type"] == SCF:
            cc.nmax, cc.lmax = c["nmax"], c["lmax"]
            cc.S = c["S"].ctypes.data
            cc.T = c["T"].ctypes.data
            keep += [c["S"], c["T"]]
    P._keep = keep
    return P


# --------------------------------------------------------------------------
# numpy-level wrappers

def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def value(components, q, t=0.0, ref=False):
    """q (3,N) -> (N,)"""
    P = make_pot(components)
    q = _c(q).reshape(3, -1)
    L = ref_lib() if ref else lib()
    f = L.ref_value if ref else L.orc_value
    return np.array([f(C.byref(P), t, _c(q[:, i])) for i in range(q.shape[1])])


def density(components, q, t=0.0, ref=False):
    P = make_pot(components)
    q = _c(q).reshape(3, -1)
    L = ref_lib() if ref else lib()
    f = L.ref_density if ref else L.orc_density
    return np.array([f(C.byref(P), t, _c(q[:, i])) for i in range(q.shape[1])])


def gradient(components, q, t=0.0, ref=False):
    """q (3,N) -> (3,N)"""
    P = make_pot(components)
    q = _c(q).reshape(3, -1)
    out = np.zeros_like(q)
    if ref:
        ref_lib().ref_gradient(C.byref(P), q.shape[1], t, q, out)
    else:
        lib().orc_gradient(C.byref(P), q.shape[1], t, q, out)
    return out


def hessian(components, q, t=0.0, ref=False):
    """q (3,N) -> (3,3,N) as gala's hessian layout"""
    P = make_pot(components)
    q = _c(q).reshape(3, -1)
    out = np.zeros((q.shape[1], 9))
    for i in range(q.shape[1]):
        h = np.zeros(9)
        if ref:
            ref_lib().ref_hessian(C.byref(P), t, _c(q[:, i]), h)
        else:
            lib().orc_hessian(C.byref(P), t, _c(q[:, i]), h)
        out[i] = h
    return out.reshape(-1, 3, 3).transpose(1, 2, 0)


def d2_dr2(components, q, t=0.0, ref=False):
    P = make_pot(components)
    q = _c(q).reshape(3, -1)
    L = ref_lib() if ref else lib()
    f = L.ref_d2_dr2 if ref else L.orc_d2_dr2
    return np.array([f(C.byref(P), t, _c(q[:, i])) for i in range(q.shape[1])])


def leapfrog(components, w0, t, save_all=True, fast=False):
    """w0 (6,N), t (ntimes,) -> (6,ntimes,N) or (6,N)"""
    P = make_pot(components)
    w0 = _c(w0).reshape(6, -1)
    t = _c(t)
    N = w0.shape[1]
    out = np.zeros((6, len(t), N) if save_all else (6, N))
    lib(fast).orc_leapfrog(C.byref(P), w0, N, t, len(t), int(save_all), out.reshape(-1))
    return out


def ruth4(components, w0, t, save_all=True, fast=False):
    """Ruth4 (ruth4.pyx:24-106): w0 (6,N), t (ntimes,) -> (6,ntimes,N) or (6,N)"""
    P = make_pot(components)
    w0 = _c(w0).reshape(6, -1)
    t = _c(t)
    N = w0.shape[1]
    out = np.zeros((6, len(t), N) if save_all else (6, N))
    lib(fast).orc_ruth4(C.byref(P), w0, N, t, len(t), int(save_all), out.reshape(-1))
    return out


def dop853(components, w0, t, atol=1e-10, rtol=1e-10, nmax=0, nstiff=0, dt_max=0.0,
           save_all=True, nbatch=100, ref=False):
    """dop853_integrate_hamiltonian: w0 (6,N), t -> (out, status[, stats])."""
    P = make_pot(components)
    w0 = _c(w0).reshape(6, -1)
    t = _c(t)
    N = w0.shape[1]
    out = np.zeros((6, len(t), N) if save_all else (6, N))
    status = np.zeros(N, dtype=np.int32)
    if ref:
        ref_lib().ref_dop853(C.byref(P), w0, N, t, len(t), atol, rtol, nmax, nstiff, dt_max,
                             int(save_all), nbatch, out.reshape(-1), status)
        return out, status
    stats = np.zeros((N, 4), dtype=np.int64)
    lib().orc_dop853(C.byref(P), w0, N, t, len(t), atol, rtol, nmax, nstiff, dt_max,
                     int(save_all), nbatch, out.reshape(-1), status, stats)
    return out, status, stats


def dop853_orbits(components, w0, t0, t1, h0, atol, rtol, nmax=0, nstiff=0, fast=False):
    """Independent per-orbit DOP853 to per-orbit end times; returns
    (w_final (6,N), status (N,), stats (N,4) = nstep, naccpt, nrejct, nfcn)."""
    P = make_pot(components)
    w0 = _c(w0).reshape(6, -1)
    N = w0.shape[1]
    t0 = _c(np.broadcast_to(t0, (N,)))
    t1 = _c(np.broadcast_to(t1, (N,)))
    h0 = _c(np.broadcast_to(h0, (N,)))
    out = np.zeros((6, N))
    status = np.zeros(N, dtype=np.int32)
    stats = np.zeros((N, 4), dtype=np.int64)
    lib(fast).orc_dop853_orbits(C.byref(P), w0, N, t0, t1, h0, atol, rtol, nmax, nstiff,
                                out.reshape(-1), status, stats)
    return out, status, stats


def _bodies(bodies):
    """Body potentials for the N-body functions: a list with one component
    list per body (None or [] = massless).  A bare component list (or None)
    is the progenitor alone.  Returns (ctypes array of orc_pot*, count,
    keepalive)."""
    if bodies is None or (len(bodies) > 0 and isinstance(bodies[0], dict)):
        bodies = [bodies]
    pots = [make_pot(b) if b else None for b in bodies]
    arr = (C.c_void_p * max(len(pots), 1))()
    for j, p in enumerate(pots):
        arr[j] = C.addressof(p) if p is not None else None
    return arr, len(pots), pots


def nbody_leapfrog(ext_components, bodies, w0, t, save_all=True, ref=False):
    """leapfrog_integrate_nbody (leapfrog.pyx:161-257).  w0 (n, 6) rows,
    bodies: potentials of bodies 0..K-1 (see _bodies).  Returns (ntimes, n, 6)
    when save_all, else (n, 6)."""
    P = make_pot(ext_components)
    arr, K, keep = _bodies(bodies)
    w0 = _c(w0).reshape(-1, 6)
    n = w0.shape[0]
    t = _c(t)
    out = np.zeros((len(t), n, 6) if save_all else (n, 6))
    L = ref_lib() if ref else lib()
    fn = L.ref_nbody_leapfrog if ref else L.orc_nbody_leapfrog
    fn(C.byref(P), arr, K, w0, n, t, len(t), int(save_all), out.reshape(-1))
    return out


def nbody_ruth4(ext_components, bodies, w0, t, save_all=True, ref=False):
    """ruth4_integrate_nbody (ruth4.pyx:138-261); layouts as nbody_leapfrog."""
    P = make_pot(ext_components)
    arr, K, keep = _bodies(bodies)
    w0 = _c(w0).reshape(-1, 6)
    n = w0.shape[0]
    t = _c(t)
    out = np.zeros((len(t), n, 6) if save_all else (n, 6))
    fn = ref_lib().ref_nbody_ruth4 if ref else lib().orc_nbody_ruth4
    fn(C.byref(P), arr, K, w0, n, t, len(t), int(save_all), out.reshape(-1))
    return out


def nbody_dop853(ext_components, bodies, w0, t, atol=1e-10, rtol=1e-10, nmax=0, dt_max=0.0,
                 save_all=True, ref=False):
    """direct_nbody_dop853 (nbody.pyx:30-115).  w0 (n, 6).  Returns (out
    ((ntimes, n, 6) or (n, 6)), idid, stats (4,) or None with ref)."""
    P = make_pot(ext_components)
    arr, K, keep = _bodies(bodies)
    w0 = _c(w0).reshape(-1, 6)
    n = w0.shape[0]
    t = _c(t)
    out = np.zeros((len(t), n, 6) if save_all else (n, 6))
    if ref:
        idid = ref_lib().ref_nbody_dop853(C.byref(P), arr, K, w0, n, t, len(t), atol, rtol, nmax,
                                          dt_max, int(save_all), out.reshape(-1))
        return out, idid, None
    stats = np.zeros(4, dtype=np.int64)
    idid = lib().orc_nbody_dop853(C.byref(P), arr, K, w0, n, t, len(t), atol, rtol, nmax,
                                  dt_max, int(save_all), out.reshape(-1), stats)
    return out, idid, stats


def mockstream_leapfrog(ext_components, bodies, full_time, body_w0, spawn_t, nstream,
                        stream_w0, tfinal, fast=False):
    """mockstream_leapfrog (mockstream.pyx:442-620).  bodies: the progenitor's
    components (or None), or a list of per-body component lists (progenitor
    first, see _bodies); body_w0 (K, 6) at full_time[0].  Returns (body_final
    ((6,) for one body, else (K, 6)), stream_final (ntot, 6))."""
    P = make_pot(ext_components)
    arr, K, keep = _bodies(bodies)
    full_time = _c(full_time)
    spawn_t = _c(spawn_t)
    nstream = np.ascontiguousarray(nstream, dtype=np.int32)
    stream_w0 = _c(stream_w0).reshape(-1, 6)
    body_final = np.zeros((K, 6))
    stream_final = np.zeros_like(stream_w0)
    lib(fast).orc_mockstream_leapfrog(C.byref(P), arr, K, full_time, len(full_time),
                                      _c(body_w0).reshape(-1), len(spawn_t), spawn_t, nstream,
                                      stream_w0, float(tfinal), body_final.reshape(-1),
                                      stream_final.reshape(-1))
    return (body_final[0] if K == 1 else body_final), stream_final


def mockstream_dop853(ext_components, bodies, body_w, stream_t1, nstream, stream_w0,
                      tfinal, dt0, atol=1e-10, rtol=1e-10, nmax=0, ref=False):
    """Release systems of mockstream_dop853 (mockstream.pyx:286-303).
    bodies as in mockstream_leapfrog; body_w (nrel, K, 6) (or (nrel, 6) for
    one body) the bodies' states at the release times; stream_t1 (nrel,);
    nstream (nrel,) int; stream_w0 (ntot, 6).  Returns (out (ntot, 6),
    body_final ((6,) for one body, else (K, 6)), status (nrel,), stats
    (nrel, 4) or None with ref)."""
    P = make_pot(ext_components)
    arr, K, keep = _bodies(bodies)
    body_w = _c(body_w).reshape(-1, K * 6)
    nrel = body_w.shape[0]
    t1 = _c(stream_t1)
    ns = np.ascontiguousarray(nstream, dtype=np.int32)
    sw = _c(stream_w0).reshape(-1, 6)
    out = np.zeros_like(sw)
    bf = np.zeros((K, 6))
    status = np.zeros(nrel, dtype=np.int32)
    one = (lambda x: x[0]) if K == 1 else (lambda x: x)
    if ref:
        ref_lib().ref_mockstream_dop853(C.byref(P), arr, K, body_w, nrel, t1, ns, sw, tfinal, dt0,
                                        atol, rtol, nmax, out, bf.reshape(-1), status)
        return out, one(bf), status, None
    stats = np.zeros((nrel, 4), dtype=np.int64)
    lib().orc_mockstream_dop853(C.byref(P), arr, K, body_w, nrel, t1, ns, sw, tfinal, dt0, atol,
                                rtol, nmax, out, bf.reshape(-1), status, stats)
    return out, one(bf), status, stats


def fardal_sample(components, G, prog_x, prog_v, prog_t, prog_m, nparticles, z, lead=True,
                  trail=True, gala_modified=True):
    """FardalStreamDF._sample with pre-drawn standard normals z."""
    P = make_pot(components)
    prog_x = _c(prog_x).reshape(-1, 3)
    prog_v = _c(prog_v).reshape(-1, 3)
    prog_t = _c(prog_t)
    prog_m = _c(np.broadcast_to(prog_m, prog_t.shape))
    nparticles = np.ascontiguousarray(nparticles, dtype=np.int32)
    ntot = int((int(lead) + int(trail)) * nparticles.sum())
    ox = np.zeros((ntot, 3))
    ov = np.zeros((ntot, 3))
    ot = np.zeros(ntot)
    lib().orc_fardal_sample(C.byref(P), G, prog_x, prog_v, prog_t, prog_m, nparticles,
                            len(prog_t), int(lead), int(trail), int(gala_modified), _c(z),
                            ox, ov, ot)
    return ox, ov, ot


def num_threads(fast=True):
    return lib(fast).orc_num_threads()
