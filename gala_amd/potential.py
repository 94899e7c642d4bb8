"""Potentials, composite potentials and the Hamiltonian: gala's Python API
for the hot path, backed by the HIP kernels of libgala_amd.so.

Mirrors (paths relative to gala's src/gala/):
  PotentialBase            potential/potential/core.py:27 (energy :356-423,
                           gradient :425-479, density, hessian,
                           circular_velocity :725-785, integrate_orbit :1150-1184)
  HernquistPotential etc.  potential/potential/builtin/core.py:215-272, 528-563, 688-741
  CompositePotential       potential/potential/core.py:1274-1508
  SCFPotential             potential/scf/bfe_class.pyx:49-106
  Hamiltonian              potential/hamiltonian/chamiltonian.pyx:40-380
  StaticFrame              potential/frame/builtin/frames.pyx:37-50

Inputs are plain arrays in the potential's unit system (kpc, Myr, Msun by
default): positions shaped (3, ...) as in gala.  Results are numpy arrays.
Every evaluation runs on the GPU through the C ABI.
"""
from collections import OrderedDict

import numpy as np

from . import _device as dev
from . import _lib
from .units import UnitSystem, galactic

__all__ = ["PotentialBase", "HernquistPotential", "PlummerPotential",
           "MiyamotoNagaiPotential", "NFWPotential", "NullPotential", "CompositePotential",
           "SCFPotential", "Hamiltonian", "StaticFrame", "MilkyWayLikePotential"]


def _as_units(units):
    if units is None:
        return galactic
    if isinstance(units, UnitSystem):
        return units
    if units == "galactic":
        return galactic
    raise TypeError(f"unsupported unit system {units!r}")


class _Handle:
    """Owns one gamd_potential (device copy of the component table)."""

    def __init__(self, components):
        lib = _lib.load()
        dev.require_gpu()
        arr, self._keep = _lib.make_components(components)
        h = _lib.C.c_void_p()
        _lib.check(lib.gamd_potential_create(arr, len(components), _lib.C.byref(h)),
                   "gamd_potential_create")
        self.ptr = h.value
        self._lib = lib

    def __del__(self):
        try:
            if self.ptr:
                self._lib.gamd_potential_destroy(self.ptr)
        except Exception:
            pass


class PotentialBase:
    """Base of all potentials (gala potential/potential/core.py:27)."""

    ndim = 3
    _type = _lib.NULL
    _param_names = ()
    _param_defaults = {}

    def __init__(self, *args, units=None, origin=None, R=None, **kwargs):
        self.units = _as_units(units)
        self.G = self.units.G
        names = self._param_names
        if len(args) > len(names):
            raise TypeError(f"{type(self).__name__} takes at most {len(names)} parameters")
        params = OrderedDict()
        for name, val in zip(names, args):
            params[name] = val
        for name in names:
            if name in kwargs:
                if name in params:
                    raise TypeError(f"parameter {name} given twice")
                params[name] = kwargs.pop(name)
            elif name not in params:
                if name in self._param_defaults:
                    params[name] = self._param_defaults[name]
                else:
                    raise TypeError(f"missing parameter {name!r} for {type(self).__name__}")
        if kwargs:
            raise TypeError(f"unexpected arguments {sorted(kwargs)}")
        self.parameters = OrderedDict((k, float(v)) for k, v in params.items())
        self.origin = np.zeros(3) if origin is None else np.asarray(origin, dtype=np.float64)
        if R is None:
            self.R = np.eye(3)
        else:
            R = np.asarray(R, dtype=np.float64)
            if R.shape != (3, 3):
                raise ValueError("R must be a 3x3 rotation matrix")
            self.R = R
        self._handle = None

    # ---- the component table the C ABI consumes --------------------------
    def _c_pars(self):
        """[G] + parameters in declaration order (cybuiltin.pyx:168)."""
        return [self.G] + [self.parameters[n] for n in self._param_names]

    def _components(self):
        shift = not (np.all(self.origin == 0.0) and np.all(self.R == np.eye(3)))
        return [{"type": self._type, "pars": self._c_pars(), "q0": self.origin,
                 "R": self.R, "shift_rotate": int(shift)}]

    @property
    def c_handle(self):
        if self._handle is None:
            self._handle = _Handle(self._components())
        return self._handle.ptr

    # ---- evaluation ---------------------------------------------------------
    def _prep_q(self, q):
        q = np.asarray(q, dtype=np.float64)
        if q.shape[0] != 3:
            raise ValueError(f"positions must have shape (3, ...), got {q.shape}")
        shape = q.shape[1:]
        return q.reshape(3, -1), shape

    def _eval(self, fn, q, t, nout):
        lib = _lib.load()
        q2, shape = self._prep_q(q)
        N = q2.shape[1]
        dq = dev.to_dev(q2)
        out = dev.empty((nout, N) if nout > 1 else (N,))
        _lib.check(getattr(lib, fn)(self.c_handle, float(t), dq.data_ptr(), N, out.data_ptr(),
                                    dev.stream_ptr()), fn)
        return dev.to_host(out), shape

    def energy(self, q, t=0.0):
        """Potential energy Phi(q); q (3, ...) -> (...)."""
        v, shape = self._eval("gamd_potential_value", q, t, 1)
        return v.reshape(shape)

    def gradient(self, q, t=0.0):
        """grad Phi; q (3, ...) -> (3, ...)."""
        g, shape = self._eval("gamd_potential_gradient", q, t, 3)
        return g.reshape((3,) + shape)

    def acceleration(self, q, t=0.0):
        return -self.gradient(q, t=t)

    def density(self, q, t=0.0):
        v, shape = self._eval("gamd_potential_density", q, t, 1)
        return v.reshape(shape)

    def hessian(self, q, t=0.0):
        """Second derivatives; q (3, ...) -> (3, 3, ...).  As gala, a rotated
        component's Hessian is not rotated back (cpotential.cpp:309-311)."""
        if any(c["shift_rotate"] and np.any(c["R"] != np.eye(3)) for c in self._components()):
            raise NotImplementedError("Hessians of rotated potentials are not supported "
                                      "(as in gala)")
        h, shape = self._eval("gamd_potential_hessian", q, t, 9)
        return h.reshape((3, 3) + shape)

    def d2_dr2(self, q, t=0.0):
        """Radial second difference of Phi (c_d2_dr2, cpotential.cpp:346-371)."""
        v, shape = self._eval("gamd_potential_d2_dr2", q, t, 1)
        return v.reshape(shape)

    def circular_velocity(self, q, t=0.0):
        """sqrt(r |dPhi/dr|) with dPhi/dr = grad Phi . r_hat (core.py:772-785)."""
        q2, shape = self._prep_q(q)
        r = np.sqrt(np.sum(q2 ** 2, axis=0))
        g = self.gradient(q2, t=t)
        dPhi_dr = np.sum(g * q2 / r, axis=0)
        return np.sqrt(r * np.abs(dPhi_dr)).reshape(shape)

    def __call__(self, q):
        return self.energy(q)

    def integrate_orbit(self, *args, **kwargs):
        """Hamiltonian(self).integrate_orbit(...) (core.py:1150-1184)."""
        return Hamiltonian(self).integrate_orbit(*args, **kwargs)

    def __add__(self, other):
        if not isinstance(other, PotentialBase):
            return NotImplemented
        comp = CompositePotential()
        for i, p in enumerate((self, other)):
            if isinstance(p, CompositePotential):
                for k, v in p.items():
                    comp[f"{k}"] = v
            else:
                comp[str(len(comp))] = p
        return comp

    def __repr__(self):
        pars = ", ".join(f"{k}={v:g}" for k, v in self.parameters.items())
        return f"<{type(self).__name__}: {pars} ({self.units})>"


class HernquistPotential(PotentialBase):
    """Hernquist sphere, Phi = -G m / (r + c) (builtin/core.py:215-242)."""
    _type = _lib.HERNQUIST
    _param_names = ("m", "c")


class PlummerPotential(PotentialBase):
    """Plummer sphere, Phi = -G m / sqrt(r^2 + b^2) (builtin/core.py:246-272)."""
    _type = _lib.PLUMMER
    _param_names = ("m", "b")


class MiyamotoNagaiPotential(PotentialBase):
    """Miyamoto-Nagai disk (builtin/core.py:528-563)."""
    _type = _lib.MIYAMOTONAGAI
    _param_names = ("m", "a", "b")


class NFWPotential(PotentialBase):
    """NFW halo (builtin/core.py:688-741).  The spherical case (a = b = c = 1)
    selects SphericalNFWWrapper there; the flattened and triaxial variants are
    outside this engine and raise."""
    _type = _lib.NFW
    _param_names = ("m", "r_s", "a", "b", "c")
    _param_defaults = {"a": 1.0, "b": 1.0, "c": 1.0}

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        p = self.parameters
        if not np.allclose([p["a"], p["b"], p["c"]], 1.0):
            raise NotImplementedError("only the spherical NFW (a = b = c = 1) is supported")


class NullPotential(PotentialBase):
    """Phi = 0 (NullWrapper, cybuiltin.pyx:362-372)."""
    _type = _lib.NULL
    _param_names = ()

    def _c_pars(self):
        return [self.G]


class SCFPotential(PotentialBase):
    """Hernquist-Ostriker basis-function expansion with Lowing+2011 real
    coefficients Snlm (cos) and Tnlm (sin), shape (nmax+1, lmax+1, lmax+1)
    (scf/bfe_class.pyx:49-106).  Terms with m > l are ignored."""
    _type = _lib.SCF
    _param_names = ("m", "r_s")

    def __init__(self, m, r_s, Snlm, Tnlm=None, units=None, origin=None, R=None):
        super().__init__(m, r_s, units=units, origin=origin, R=R)
        S = np.ascontiguousarray(Snlm, dtype=np.float64)
        T = np.zeros_like(S) if Tnlm is None else np.ascontiguousarray(Tnlm, dtype=np.float64)
        if S.ndim != 3 or S.shape != T.shape or S.shape[1] != S.shape[2]:
            raise ValueError("Snlm and Tnlm must both have shape (nmax+1, lmax+1, lmax+1)")
        nmax, lmax = S.shape[0] - 1, S.shape[1] - 1
        if nmax > _lib.SCF_MAX_NMAX or lmax > _lib.SCF_MAX_LMAX:
            raise ValueError(f"nmax <= {_lib.SCF_MAX_NMAX} and lmax <= {_lib.SCF_MAX_LMAX}")
        self.Snlm, self.Tnlm = S, T
        self.nmax, self.lmax = nmax, lmax

    def _components(self):
        c = super()._components()[0]
        c["S"], c["T"] = self.Snlm, self.Tnlm
        return [c]


class CompositePotential(PotentialBase, OrderedDict):
    """Ordered collection of potentials summed in insertion order
    (core.py:1274-1508; C layout ccompositepotential.pyx:27-70)."""

    def __init__(self, *args, **kwargs):
        units = kwargs.pop("units", None)
        OrderedDict.__init__(self)
        self.units = _as_units(units)
        self.G = self.units.G
        self.parameters = OrderedDict()
        self.origin = np.zeros(3)
        self.R = np.eye(3)
        self._handle = None
        self._locked = False
        if args and isinstance(args[0], dict):
            for k, v in args[0].items():
                self[k] = v
        for k, v in kwargs.items():
            self[k] = v

    def __setitem__(self, key, value):
        if getattr(self, "_locked", False):
            raise ValueError("potential components cannot be changed once the potential has "
                             "been used on the GPU")
        if not isinstance(value, PotentialBase):
            raise TypeError("components must be potentials")
        if value.units != self.units:
            raise ValueError("all components must share one unit system")
        OrderedDict.__setitem__(self, key, value)
        self._handle = None

    def _components(self):
        out = []
        for p in self.values():
            out += p._components()
        if len(out) > _lib.MAX_COMPONENTS:
            raise ValueError(f"at most {_lib.MAX_COMPONENTS} components")
        return out

    @property
    def c_handle(self):
        self._locked = True
        return PotentialBase.c_handle.fget(self)

    def __repr__(self):
        return "<CompositePotential " + ",".join(self.keys()) + ">"

    def __hash__(self):
        return id(self)

    def __eq__(self, other):
        return self is other


def MilkyWayLikePotential(units=None):
    """The four-component Milky-Way-like composite of the benchmark
    workloads: Hernquist nucleus (1.7e9 Msun, 0.07 kpc), Hernquist bulge
    (5e9 Msun, 1 kpc), Miyamoto-Nagai disk (6.8e10 Msun, 3 kpc, 0.28 kpc) and
    NFW halo (5.4e11 Msun, 15.6 kpc), the layout of gala's MilkyWayPotential
    (potential/potential/builtin/special.py:98-124)."""
    pot = CompositePotential(units=units)
    pot["nucleus"] = HernquistPotential(m=1.7e9, c=0.07, units=units)
    pot["bulge"] = HernquistPotential(m=5e9, c=1.0, units=units)
    pot["disk"] = MiyamotoNagaiPotential(m=6.8e10, a=3.0, b=0.28, units=units)
    pot["halo"] = NFWPotential(m=5.4e11, r_s=15.6, units=units)
    return pot


class StaticFrame:
    """Static, inertial frame: H = Phi(q) + |p|^2/2 (builtin_frames.cpp:8-24)."""

    def __init__(self, units=None):
        self.units = _as_units(units)

    def __eq__(self, other):
        return isinstance(other, StaticFrame) and other.units == self.units

    def __hash__(self):
        return hash(("StaticFrame", self.units))


class Hamiltonian:
    """Potential + reference frame (chamiltonian.pyx:40-77).  Only the static
    frame is supported (the rotating frames are outside this engine)."""

    def __init__(self, potential, frame=None):
        if isinstance(potential, Hamiltonian):
            frame = potential.frame
            potential = potential.potential
        if frame is None:
            frame = StaticFrame(potential.units)
        if not isinstance(frame, StaticFrame):
            raise NotImplementedError("only StaticFrame is supported")
        self.potential = potential
        self.frame = frame
        self.units = potential.units
        self.c_enabled = True

    @property
    def ndim(self):
        return 2 * self.potential.ndim

    def energy(self, w, t=0.0):
        """H(w) = Phi(q) + |p|^2/2; w (6, ...) -> (...)."""
        w = np.asarray(w, dtype=np.float64)
        if w.shape[0] != 6:
            raise ValueError("w must have shape (6, ...)")
        shape = w.shape[1:]
        w2 = w.reshape(6, -1)
        lib = _lib.load()
        dw = dev.to_dev(w2)
        out = dev.empty((w2.shape[1],))
        _lib.check(lib.gamd_hamiltonian_energy(self.potential.c_handle, float(t), dw.data_ptr(),
                                               w2.shape[1], out.data_ptr(), dev.stream_ptr()),
                   "gamd_hamiltonian_energy")
        return dev.to_host(out).reshape(shape)

    def __call__(self, w):
        return self.energy(w)

    def integrate_orbit(self, w0, Integrator=None, Integrator_kwargs=None,
                        cython_if_possible=True, save_all=True, **time_spec):
        """Integrate orbits (chamiltonian.pyx:224-380); see
        gala_amd.integrate.integrate_hamiltonian."""
        from .integrate import integrate_hamiltonian
        return integrate_hamiltonian(self, w0, Integrator=Integrator,
                                     Integrator_kwargs=Integrator_kwargs or {},
                                     save_all=save_all, **time_spec)
