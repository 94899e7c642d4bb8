"""Phase-space containers, orbits and mock stellar streams.

Mirrors (gala src/gala/...):
  PhaseSpacePosition     dynamics/core.py (from_w :528-567)
  Orbit                  dynamics/orbit.py:22-389 (layout :74-175, energy :361-389)
  MockStream             dynamics/mockstream/core.py
  FardalStreamDF         dynamics/mockstream/df.pyx:125-238 (sample), 363-456
                         (_sample), 61-106 (get_rj_vj_R, transform_from_sat)
  MockStreamGenerator    dynamics/mockstream/mockstream_generator.py:19-372
  DirectNBody            dynamics/nbody/core.py (the progenitor-only case)

Arrays are in the unit system of the potential (kpc, Myr, Msun by default):
positions (3, ...), velocities (3, ...) in kpc/Myr, times in Myr.  Only the
static frame exists here, so frame transformations are identities.
"""
import numpy as np

from . import _device as dev
from . import _lib
from .integrate import (DOPRI853Integrator, LeapfrogIntegrator, get_integrator, leapfrog,
                        parse_time_specification)
from .potential import Hamiltonian, NullPotential, PotentialBase, StaticFrame

__all__ = ["PhaseSpacePosition", "Orbit", "MockStream", "FardalStreamDF", "StreaklineStreamDF",
           "MockStreamGenerator", "DirectNBody"]


class PhaseSpacePosition:
    """Positions and velocities, each shaped (3, ...) (dynamics/core.py)."""

    def __init__(self, pos, vel, frame=None):
        self.xyz = np.asarray(pos, dtype=np.float64)
        self.v_xyz = np.asarray(vel, dtype=np.float64)
        if self.xyz.shape[0] != 3 or self.v_xyz.shape != self.xyz.shape:
            raise ValueError("pos and vel must both have shape (3, ...)")
        self.frame = frame

    pos = property(lambda self: self.xyz)
    vel = property(lambda self: self.v_xyz)
    ndim = 3

    @property
    def shape(self):
        return self.xyz.shape[1:]

    def w(self, units=None):
        """(6, ...) stack of positions and velocities."""
        return np.vstack((self.xyz[None], self.v_xyz[None])).reshape((6,) + self.shape)

    @classmethod
    def from_w(cls, w, units=None, frame=None):
        w = np.asarray(w)
        return cls(pos=w[:3], vel=w[3:], frame=frame)

    def energy(self, hamiltonian):
        H = Hamiltonian(hamiltonian)
        return H.energy(self.w())

    def kinetic_energy(self):
        return 0.5 * np.sum(self.v_xyz ** 2, axis=0)

    def potential_energy(self, potential):
        return potential.energy(self.xyz)

    def __getitem__(self, idx):
        if not isinstance(idx, tuple):
            idx = (idx,)
        return type(self)(self.xyz[(slice(None),) + idx], self.v_xyz[(slice(None),) + idx],
                          frame=self.frame)

    def __len__(self):
        return self.shape[0] if self.shape else 0


class Orbit(PhaseSpacePosition):
    """A set of orbits: t (ntimes,), pos/vel (3, ntimes[, norbits])
    (dynamics/orbit.py:74-106)."""

    def __init__(self, pos, vel, t=None, hamiltonian=None, potential=None, frame=None):
        super().__init__(pos, vel, frame=frame)
        if t is not None:
            t = np.atleast_1d(np.asarray(t, dtype=np.float64))
            if self.xyz.ndim < 2 or t.shape[0] != self.xyz.shape[1]:
                raise ValueError("t must match the time axis (axis 1) of pos")
        self.t = t
        if hamiltonian is not None:
            self.hamiltonian = Hamiltonian(hamiltonian)
            self.potential = self.hamiltonian.potential
            self.frame = self.hamiltonian.frame
        else:
            self.hamiltonian = Hamiltonian(potential, frame) if potential is not None else None
            self.potential = potential

    @classmethod
    def from_w(cls, w, units=None, t=None, hamiltonian=None, **kw):
        w = np.asarray(w)
        return cls(pos=w[:3], vel=w[3:], t=t, hamiltonian=hamiltonian)

    @property
    def ntimes(self):
        return self.xyz.shape[1]

    @property
    def norbits(self):
        return 1 if self.xyz.ndim == 2 else self.xyz.shape[2]

    def energy(self, hamiltonian=None):
        """Total energy per unit mass (orbit.py:361-389)."""
        if self.hamiltonian is None and hamiltonian is None:
            raise ValueError("To compute the total energy, a hamiltonian object must be provided!")
        H = self.hamiltonian if hamiltonian is None else Hamiltonian(hamiltonian)
        return H.energy(self.w())

    def __getitem__(self, idx):
        if not isinstance(idx, tuple):
            idx = (idx,)
        t = self.t
        if t is not None and len(idx) >= 1:
            t = t[idx[0]]
        pos = self.xyz[(slice(None),) + idx]
        vel = self.v_xyz[(slice(None),) + idx]
        if t is None or np.ndim(t) == 0:
            return PhaseSpacePosition(pos, vel, frame=self.frame)
        return Orbit(pos, vel, t=t, hamiltonian=self.hamiltonian)


class MockStream(PhaseSpacePosition):
    """Stream particles with release times and 't'/'l' tail labels
    (dynamics/mockstream/core.py)."""

    def __init__(self, pos, vel, release_time=None, lead_trail=None, frame=None):
        super().__init__(pos, vel, frame=frame)
        self.release_time = None if release_time is None else np.asarray(release_time)
        self.lead_trail = None if lead_trail is None else np.asarray(lead_trail)


# ---------------------------------------------------------------------------
# stream distribution functions

class BaseStreamDF:
    """Release geometry shared by the stream DFs (df.pyx:32-123)."""

    def __init__(self, lead=True, trail=True, random_state=None):
        self._lead = int(lead)
        self._trail = int(trail)
        if random_state is None:
            random_state = np.random.RandomState()
        self.random_state = random_state
        if not self.lead and not self.trail:
            raise ValueError("You must generate either leading or trailing tails (or both!)")

    lead = property(lambda self: self._lead)
    trail = property(lambda self: self._trail)

    @staticmethod
    def _rj_vj_R(potential, G, prog_x, prog_v, prog_m, prog_t):
        """get_rj_vj_R (df.pyx:61-92) for every release time at once.
        prog_x, prog_v (ntimes, 3).  Returns rj, vj (ntimes,), R (ntimes, 3, 3)
        (rows: x_hat, -(x_hat x L_hat), L_hat)."""
        x0, x1, x2 = prog_x[:, 0], prog_x[:, 1], prog_x[:, 2]
        v0, v1, v2 = prog_v[:, 0], prog_v[:, 1], prog_v[:, 2]
        dist = np.sqrt(x0 * x0 + x1 * x1 + x2 * x2)
        L = np.stack([x1 * v2 - x2 * v1, -x0 * v2 + x2 * v0, x0 * v1 - x1 * v0], axis=1)
        Lnorm = np.sqrt(L[:, 0] * L[:, 0] + L[:, 1] * L[:, 1] + L[:, 2] * L[:, 2])
        R = np.empty((len(dist), 3, 3))
        R[:, 0, :] = prog_x / dist[:, None]
        R[:, 2, :] = L / Lnorm[:, None]
        Om = Lnorm / (dist * dist)
        # c_d2_dr2 of the external potential at every release point, one launch
        d2r = potential.d2_dr2(np.ascontiguousarray(prog_x.T)) if len(dist) else np.zeros(0)
        rj = (G * prog_m / (Om * Om - d2r)) ** (1 / 3.0)
        vj = Om * rj
        a, b = R[:, 0, :], R[:, 2, :]
        R[:, 1, 0] = -(a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1])
        R[:, 1, 1] = -(-a[:, 0] * b[:, 2] + a[:, 2] * b[:, 0])
        R[:, 1, 2] = -(a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0])
        return rj, vj, R

    @staticmethod
    def _from_sat(R, tx, tv, px, pv):
        """transform_from_sat (df.pyx:94-106): R^T x + prog, per particle."""
        out_x = (R[:, 0, :] * tx[:, 0:1] + R[:, 1, :] * tx[:, 1:2]) + R[:, 2, :] * tx[:, 2:3] + px
        out_v = (R[:, 0, :] * tv[:, 0:1] + R[:, 1, :] * tv[:, 1:2]) + R[:, 2, :] * tv[:, 2:3] + pv
        return out_x, out_v

    def sample(self, prog_orbit, prog_mass, hamiltonian=None, release_every=1, n_particles=1):
        """Stream-particle initial conditions and release times (df.pyx:125-238)."""
        if getattr(prog_orbit, "hamiltonian", None) is not None:
            H = prog_orbit.hamiltonian
        elif hamiltonian is not None:
            H = Hamiltonian(hamiltonian)
        else:
            raise ValueError("a hamiltonian is needed to sample the stream DF")
        prog_x = np.ascontiguousarray(prog_orbit.xyz.T)
        prog_v = np.ascontiguousarray(prog_orbit.v_xyz.T)
        prog_t = np.asarray(prog_orbit.t, dtype=np.float64)
        prog_m = np.squeeze(np.asarray(prog_mass, dtype=np.float64))
        if prog_m.shape == ():
            prog_m = np.full_like(prog_t, prog_m)
        if np.iterable(n_particles):
            n_particles = np.array(n_particles).astype("i4")
            if not len(n_particles) == len(prog_t):
                raise ValueError("If passing in an array n_particles, its shape must match the "
                                 "number of timesteps in the progenitor orbit.")
        else:
            N = int(n_particles)
            n_particles = np.zeros_like(prog_t, dtype="i4")
            n_particles[::release_every] = N
        x, v, t1 = self._sample(H.potential, prog_x, prog_v, prog_t, prog_m, n_particles)
        # per release time a trailing block then a leading block (the
        # reference's loop over n_particles, whose slices past the end of the
        # sampled particles are empty)
        labels = np.array([c for c, on in (("t", self._trail), ("l", self._lead)) if on],
                          dtype="U1")
        lt = np.repeat(np.tile(labels, len(n_particles)),
                       np.repeat(np.asarray(n_particles, dtype=np.int64), len(labels)))[:len(t1)]
        return MockStream(pos=x.T, vel=v.T, release_time=t1, lead_trail=lt,
                          frame=StaticFrame(H.units))


class FardalStreamDF(BaseStreamDF):
    """Fardal+2015 particle release (df.pyx:314-456), gala_modified
    parameters by default (kx ~ N(2, 0.5), z ~ N(0, 0.5), vt ~ N(0.3, 0.5) kx,
    vz ~ N(0, 0.5) in units of r_J, v_J)."""

    def __init__(self, gala_modified=None, lead=True, trail=True, random_state=None):
        super().__init__(lead=lead, trail=trail, random_state=random_state)
        if gala_modified is None:
            gala_modified = True
        self._gala_modified = int(gala_modified)

    def _draws(self, n):
        """n standard normals, in the order the reference's successive
        random_state.normal(mean, sd) calls consume them; normal(mean, sd) is
        mean + sd * z for both RandomState and Generator."""
        return np.asarray(self.random_state.standard_normal(n), dtype=np.float64)

    def _sample(self, potential, prog_x, prog_v, prog_t, prog_m, nparticles):
        kd0 = 0.5 if self._gala_modified else 0.4
        kd4 = 0.5 if self._gala_modified else 0.4
        use = prog_m != 0
        idx = np.nonzero(use)[0]
        tails = [s for s, on in ((1.0, self._trail), (-1.0, self._lead)) if on]
        rj, vj, R = self._rj_vj_R(potential, potential.G, prog_x[idx], prog_v[idx], prog_m[idx],
                                  prog_t[idx])
        counts = np.asarray(nparticles)[idx]
        # particle order: per release time, trailing block then leading block
        rel = np.repeat(np.arange(len(idx)), counts * len(tails))
        sign = np.repeat(np.tile(np.array(tails), len(counts)), np.repeat(counts, len(tails)))
        n = len(rel)
        z = self._draws(4 * n).reshape(n, 4)
        kx = 2.0 + kd0 * z[:, 0]
        srj = sign * rj[rel]
        svj = sign * vj[rel]
        tx = np.zeros((n, 3))
        tv = np.zeros((n, 3))
        tx[:, 0] = kx * srj
        tx[:, 2] = (0.0 + 0.5 * z[:, 1]) * srj
        tv[:, 1] = (0.3 + kd4 * z[:, 2]) * svj
        if self._gala_modified:
            tv[:, 1] *= kx
        tv[:, 2] = (0.0 + 0.5 * z[:, 3]) * svj
        x, v = self._from_sat(R[rel], tx, tv, prog_x[idx][rel], prog_v[idx][rel])
        return x, v, prog_t[idx][rel]


class StreaklineStreamDF(BaseStreamDF):
    """Streakline release at +-r_J with +-v_J (df.pyx:240-311)."""

    def _sample(self, potential, prog_x, prog_v, prog_t, prog_m, nparticles):
        idx = np.nonzero(prog_m != 0)[0]
        tails = [s for s, on in ((1.0, self._trail), (-1.0, self._lead)) if on]
        rj, vj, R = self._rj_vj_R(potential, potential.G, prog_x[idx], prog_v[idx], prog_m[idx],
                                  prog_t[idx])
        counts = np.asarray(nparticles)[idx]
        rel = np.repeat(np.arange(len(idx)), counts * len(tails))
        sign = np.repeat(np.tile(np.array(tails), len(counts)), np.repeat(counts, len(tails)))
        n = len(rel)
        tx = np.zeros((n, 3))
        tv = np.zeros((n, 3))
        tx[:, 0] = sign * rj[rel]
        tv[:, 1] = sign * vj[rel]
        x, v = self._from_sat(R[rel], tx, tv, prog_x[idx][rel], prog_v[idx][rel])
        return x, v, prog_t[idx][rel]


# ---------------------------------------------------------------------------

class DirectNBody:
    """Bodies with their own potentials in an external potential
    (dynamics/nbody/core.py:22-283).  particle_potentials[i] is body i's
    potential, centred on body i (None: a massless test particle).
    integrate_orbit puts the massive bodies first, as the reference does
    (core.py:212-227), and returns the orbits in that order.  Integrators:
    DOP853 (the default; direct_nbody_dop853, nbody.pyx:30-115: all bodies one
    system with a shared adaptive step, dense output on t, h0 = t[1] - t[0],
    stiffness test off), Leapfrog (leapfrog_integrate_nbody,
    leapfrog.pyx:161-257) and Ruth4 (ruth4_integrate_nbody,
    ruth4.pyx:138-261).  Limits: at most 8 massive bodies; DOP853 takes at
    most 64 bodies in all."""

    def __init__(self, w0, particle_potentials, external_potential=None, frame=None, units=None,
                 save_all=True):
        if isinstance(w0, PhaseSpacePosition):
            w = w0.w()
        else:
            w = np.asarray(w0, dtype=np.float64)
        self.w0 = np.ascontiguousarray(w.reshape(6, -1))
        if self.w0.shape[1] != len(particle_potentials):
            raise ValueError("The number of initial conditions in `w0` must match the number of "
                             "particle potentials passed in with `particle_potentials`.")
        self.particle_potentials = [NullPotential() if p is None else p
                                    for p in particle_potentials]
        if external_potential is None:
            external_potential = NullPotential()
        self.external_potential = external_potential
        self.H = Hamiltonian(external_potential, frame)
        self.frame = self.H.frame
        self.units = self.H.units
        self.save_all = save_all

    def _reorganized(self):
        """(w0 with massive bodies first, their potentials) (core.py:212-227)."""
        massive = [i for i, p in enumerate(self.particle_potentials)
                   if not isinstance(p, NullPotential)]
        test = [i for i, p in enumerate(self.particle_potentials) if isinstance(p, NullPotential)]
        if len(massive) > _lib.MAX_MASSIVE:
            raise NotImplementedError(f"at most {_lib.MAX_MASSIVE} massive bodies are supported")
        w0 = np.ascontiguousarray(self.w0[:, massive + test])
        return w0, [self.particle_potentials[i] for i in massive]

    def integrate_orbit(self, Integrator=None, Integrator_kwargs=None, **time_spec):
        from .integrate import Ruth4Integrator, nbody_dop853, nbody_leapfrog, nbody_ruth4
        Integrator = get_integrator(Integrator or DOPRI853Integrator)
        t = parse_time_specification(self.units, **time_spec)
        w0, bodies = self._reorganized()
        if Integrator is DOPRI853Integrator:
            if w0.shape[1] > _lib.MAX_BODIES:
                raise NotImplementedError(f"DOP853 N-body systems hold at most {_lib.MAX_BODIES} "
                                          "bodies")
            kw = dict(Integrator_kwargs or {})
            res = nbody_dop853(self.external_potential, bodies, w0, t[0], t[-1], t[1] - t[0],
                               rtol=float(kw.get("rtol", 1e-10)), atol=float(kw.get("atol", 1e-10)),
                               nmax=int(kw.get("nmax", 0)), nstiff=-1,
                               hmax=float(kw.get("dt_max", 0.0)),
                               t_out=t if self.save_all else None)
            st = dev.to_host(res["status"])
            if int(kw.get("err_if_fail", 1)) == 1 and np.any(st < 0):
                raise RuntimeError(f"Integration failed with code {int(st[0])}")
            wout, wfin = res.get("w_out"), res["w_final"]
        elif Integrator is LeapfrogIntegrator:
            wout, wfin = nbody_leapfrog(self.external_potential, bodies, w0, t[1] - t[0],
                                        len(t) - 1, save=self.save_all)
        elif Integrator is Ruth4Integrator:
            wout, wfin = nbody_ruth4(self.external_potential, bodies, w0, t[1] - t[0],
                                     len(t) - 1, save=self.save_all)
        else:
            raise NotImplementedError("N-body integration supports DOP853, Leapfrog and Ruth4")
        if self.save_all:
            w = dev.to_host(wout)  # (6, ntimes, n)
            return Orbit(pos=w[:3], vel=w[3:], t=t, hamiltonian=self.H)
        w = dev.to_host(wfin)
        return PhaseSpacePosition(w[:3], w[3:], frame=self.frame)


def release_counts(orbit_t, t1s, counts):
    """The release bookkeeping of mockstream_generator.py:272-276: for each
    release time t1s[k] in order, every grid time with np.isclose(orbit_t,
    t1s[k]) gets counts[k] (a later k overwrites an earlier one).  Exact and
    vectorised: candidates come from a binary search of the sorted grid and
    are confirmed with np.isclose itself."""
    orbit_t = np.asarray(orbit_t, dtype=np.float64)
    t1s = np.asarray(t1s, dtype=np.float64)
    out = np.zeros(len(orbit_t), dtype=int)
    if len(orbit_t) == 0 or len(t1s) == 0:
        return out
    order = np.argsort(orbit_t, kind="stable")
    ts = orbit_t[order]
    tol = 1e-8 + 1e-5 * np.abs(t1s)  # np.isclose's default atol + rtol |b|
    lo = np.maximum(np.searchsorted(ts, t1s - 2 * tol, side="left") - 1, 0)
    hi = np.minimum(np.searchsorted(ts, t1s + 2 * tol, side="right") + 1, len(ts))
    width = np.maximum(hi - lo, 0)
    k = np.repeat(np.arange(len(t1s)), width)
    pos = np.repeat(lo - np.concatenate([[0], np.cumsum(width)[:-1]]), width) + np.arange(width.sum())
    hit = np.isclose(ts[pos], t1s[k])
    last = np.full(len(orbit_t), -1)
    np.maximum.at(last, order[pos[hit]], k[hit])
    sel = last >= 0
    out[sel] = np.asarray(counts)[last[sel]]
    return out


_SIDE_STREAMS = {}


def _side_stream():
    """One extra stream per device for the whole process (default priority):
    a different stream on every call (torch hands them out in turn from its
    pool) made the 8,002-particle stream take 44 ms end to end in a process
    that had run other work first, against 7.8 ms with this one stream."""
    t = dev.torch()
    d = t.cuda.current_device()
    s = _SIDE_STREAMS.get(d)
    if s is None:
        s = _SIDE_STREAMS[d] = t.cuda.Stream()
    return s


class MockStreamGenerator:
    """Generate a mock stream (mockstream_generator.py:19-372).  Leapfrog
    runs the release-and-integrate loop in the HIP kernel
    gamd_mockstream_leapfrog; the stream DF runs on the host with one GPU
    launch for the tidal-radius derivatives."""

    def __init__(self, df, hamiltonian, progenitor_potential=None):
        if not isinstance(df, BaseStreamDF):
            raise TypeError("The input distribution function (DF) instance must be an instance "
                            f"of a subclass of BaseStreamDF, not {type(df)}.")
        self.df = df
        self.hamiltonian = Hamiltonian(hamiltonian)
        if progenitor_potential is not None and not isinstance(progenitor_potential,
                                                               PotentialBase):
            raise TypeError("If specified, the progenitor_potential must be a potential instance.")
        self.self_gravity = progenitor_potential is not None
        self.progenitor_potential = progenitor_potential
        # when set, run() records HIP events around the release-and-integrate
        # kernel (Leapfrog) into kernel_events, for timing on the launch stream
        self.record_kernel_time = False
        self.kernel_events = None

    def _system(self, pw0, nbody):
        """The bodies integrated with the stream (_get_nbody,
        mockstream_generator.py:79-117): the progenitor (body 0, potential
        progenitor_potential) and the perturbers of `nbody`.  Returns (w0 (6, K),
        body potentials (None = massless))."""
        prog_pot = self.progenitor_potential
        if nbody is None:
            return pw0, [prog_pot]
        if nbody.external_potential != self.hamiltonian.potential:
            raise ValueError("The external potential of the input nbody instance must match the "
                             "potential of the mock stream input hamiltonian! "
                             f"{nbody.external_potential} vs. {self.hamiltonian.potential}")
        pots = [prog_pot] + [None if isinstance(p, NullPotential) else p
                             for p in nbody.particle_potentials]
        if any(p is None for p in pots):
            # DirectNBody.integrate_orbit would move massless bodies behind the
            # massive ones while the release systems keep the original order
            raise NotImplementedError("with perturbers, the progenitor and every perturber must "
                                      "have a potential")
        if len(pots) > _lib.MAX_MASSIVE:
            raise NotImplementedError(f"at most {_lib.MAX_MASSIVE} bodies (progenitor and "
                                      "perturbers) are supported")
        return np.ascontiguousarray(np.hstack([pw0, nbody.w0])), pots

    def run(self, prog_w0, prog_mass, nbody=None, release_every=1, n_particles=1,
            output_every=None, output_filename=None, check_filesize=True, overwrite=False,
            progress=False, Integrator=None, Integrator_kwargs=None, **time_spec):
        """Returns (stream MockStream, bodies PhaseSpacePosition (3, K): the
        progenitor and any perturbers at the final time)."""
        from .integrate import nbody_leapfrog
        if output_every is not None:
            raise NotImplementedError("snapshot output (output_every=) is not supported")
        Integrator = get_integrator(Integrator or DOPRI853Integrator)
        if Integrator not in (LeapfrogIntegrator, DOPRI853Integrator):
            raise NotImplementedError("mock streams are generated with Integrator='leapfrog' "
                                      "or 'dop853'")
        units = self.hamiltonian.units
        pot = self.hamiltonian.potential
        t = parse_time_specification(units, **time_spec)
        if isinstance(prog_w0, PhaseSpacePosition):
            pw0 = prog_w0.w().reshape(6, 1)
        else:
            pw0 = np.asarray(prog_w0, dtype=np.float64).reshape(6, 1)
        sys_w0, bodies = self._system(pw0, nbody)
        K = len(bodies)
        if Integrator is DOPRI853Integrator:
            return self._run_dop853(sys_w0, bodies, prog_mass, t, release_every, n_particles,
                                    dict(Integrator_kwargs or {}))
        # the bodies' orbits over t (backward if dt < 0), then reversed
        # (DirectNBody.integrate_orbit -> leapfrog_integrate_nbody)
        wout, _ = nbody_leapfrog(pot, bodies, sys_w0, t[1] - t[0], len(t) - 1, save=True)
        w = dev.to_host(wout)  # (6, ntimes, K)
        orbit_t = t
        if t[1] < t[0]:
            w = w[:, ::-1]
            orbit_t = t[::-1]
        # the bodies' forward trajectory on the full grid (mockstream.pyx:536-549),
        # on a side stream (default priority), so that it runs while the DF is
        # sampled (host work and the DF's own d2_dr2 launch)
        dt = orbit_t[1] - orbit_t[0]
        nfull = len(orbit_t)
        t_ = dev.torch()
        main = t_.cuda.current_stream()
        side = _side_stream()
        side.wait_stream(main)
        with t_.cuda.stream(side):
            traj_dev, _ = nbody_leapfrog(pot, bodies, np.ascontiguousarray(w[:, 0, :]), dt,
                                         nfull - 1, save=True)  # (6, nfull, K)
        prog_orbit = Orbit(pos=w[:3, :, 0], vel=w[3:, :, 0], t=orbit_t,
                           hamiltonian=self.hamiltonian)
        stream_w0 = self.df.sample(prog_orbit, prog_mass, hamiltonian=self.hamiltonian,
                                   release_every=release_every, n_particles=n_particles)
        # release bookkeeping (mockstream_generator.py:272-283)
        unq_t1s, nstream = np.unique(stream_w0.release_time, return_counts=True)
        all_nstream = release_counts(orbit_t, unq_t1s, nstream)
        main.wait_stream(side)
        traj_dev.record_stream(main)
        w0s = np.ascontiguousarray(stream_w0.w())
        n = w0s.shape[1]
        rel_t = np.asarray(stream_w0.release_time, dtype=np.float64)
        out = dev.empty((6, n))
        lib = _lib.load()
        w0d = dev.to_dev(w0s)
        reld = dev.to_dev(rel_t)
        tfinal = float(orbit_t[-1])
        handles = _lib.body_array([None if b is None else b.c_handle for b in bodies])
        if self.record_kernel_time:
            t_ = dev.torch()
            self.kernel_events = (t_.cuda.Event(enable_timing=True),
                                  t_.cuda.Event(enable_timing=True))
            self.kernel_events[0].record(t_.cuda.current_stream())
        _lib.check(lib.gamd_mockstream_leapfrog(pot.c_handle, handles, K, traj_dev.data_ptr(),
                                                nfull, float(orbit_t[0]), float(dt),
                                                w0d.data_ptr(), reld.data_ptr(), n, tfinal,
                                                out.data_ptr(), dev.stream_ptr()),
                   "gamd_mockstream_leapfrog")
        if self.record_kernel_time:
            self.kernel_events[1].record(dev.torch().cuda.current_stream())
        raw_stream = dev.to_host(out)
        traj = dev.to_host(traj_dev)
        # the bodies the reference returns: those of the last release batch,
        # integrated from their release states to tfinal
        nz = np.nonzero(all_nstream)[0]
        last_idx = int(nz[-1]) if len(nz) else nfull - 1
        if last_idx == nfull - 1:
            body_final = traj[:, -1, :]
        else:
            _, bf = nbody_leapfrog(pot, bodies, np.ascontiguousarray(traj[:, last_idx, :]), dt,
                                   nfull - 1 - last_idx, save=False)
            body_final = dev.to_host(bf)
        stream = MockStream(pos=raw_stream[:3], vel=raw_stream[3:],
                            release_time=stream_w0.release_time,
                            lead_trail=stream_w0.lead_trail, frame=self.hamiltonian.frame)
        nbody_w = PhaseSpacePosition(pos=body_final[:3], vel=body_final[3:],
                                     frame=self.hamiltonian.frame)
        return stream, nbody_w

    def _run_dop853(self, sys_w0, bodies, prog_mass, t, release_every, n_particles, kw):
        """The DOP853 path of run() (mockstream_generator.py:229-330 with
        mockstream_dop853, mockstream.pyx:176-303):
          1. the bodies' orbits over t by DOP853 as one system with dense
             output at t (DirectNBody.integrate_orbit -> direct_nbody_dop853,
             stiffness test off), reversed when t runs backwards;
          2. the DF samples particles along the progenitor's (body 0's) orbit;
          3. the bodies again, forward from the earliest time, densely output
             at the release times (dop853_helper in mockstream_dop853);
          4. every release's bodies + particles system by DOP853 to the final
             time (gamd_mockstream_dop853)."""
        from .integrate import nbody_dop853
        pot = self.hamiltonian.potential
        K = len(bodies)
        atol = float(kw.get("atol", 1e-10))
        rtol = float(kw.get("rtol", 1e-10))
        nmax = int(kw.get("nmax", 0))
        dt_max = float(kw.get("dt_max", 0.0))
        nstiff = int(kw.get("nstiff", -1))
        err_if_fail = int(kw.get("err_if_fail", 1))

        def checked(res, what):
            st = dev.to_host(res["status"])
            if err_if_fail == 1 and np.any(st < 0):
                raise RuntimeError(f"Integration failed with code {int(st[st < 0][0])} ({what})")
            return res

        res = checked(nbody_dop853(pot, bodies, sys_w0, t[0], t[-1], t[1] - t[0], rtol=rtol,
                                   atol=atol, nmax=nmax, nstiff=-1, hmax=dt_max, t_out=t),
                      "bodies' orbits")
        w = dev.to_host(res["w_out"])  # (6, ntimes, K)
        orbit_t = t
        if t[1] < t[0]:
            w = w[:, ::-1]
            orbit_t = t[::-1]
        prog_orbit = Orbit(pos=w[:3, :, 0], vel=w[3:, :, 0], t=orbit_t,
                           hamiltonian=self.hamiltonian)
        stream_w0 = self.df.sample(prog_orbit, prog_mass, hamiltonian=self.hamiltonian,
                                   release_every=release_every, n_particles=n_particles)
        # release bookkeeping (mockstream_generator.py:272-292)
        unq_t1s, nstream = np.unique(stream_w0.release_time, return_counts=True)
        all_nstream = release_counts(orbit_t, unq_t1s, nstream)
        nstream_idx = np.where(all_nstream != 0)[0]
        if 0 not in nstream_idx:
            nstream_idx = np.insert(nstream_idx, 0, 0)
            unq_t1s = np.insert(unq_t1s, 0, orbit_t[0])
        time = orbit_t[nstream_idx]
        ns = all_nstream[nstream_idx].astype(np.int32)
        if len(time) < 2:
            raise ValueError("a DOP853 mock stream needs at least two release times")
        if K + ns.max() > _lib.MAX_BODIES:
            raise NotImplementedError(f"a release system holds at most {_lib.MAX_BODIES} bodies "
                                      "(progenitor, perturbers and the release's particles)")
        # the bodies, forward, at the release times (mockstream.pyx:274-282)
        res2 = checked(nbody_dop853(pot, bodies, np.ascontiguousarray(w[:, 0, :]), time[0],
                                    time[-1], time[1] - time[0], rtol=rtol, atol=atol, nmax=nmax,
                                    nstiff=nstiff, hmax=dt_max, t_out=time),
                       "bodies at releases")
        body_w = dev.to_dev(np.ascontiguousarray(
            dev.to_host(res2["w_out"]).transpose(1, 2, 0)))  # (nrel, K, 6)
        nrel = len(time)
        offsets = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
        sw = np.ascontiguousarray(stream_w0.w().T)
        ntot = sw.shape[0]
        t_ = dev.torch()
        out = t_.empty((max(ntot, 1), 6), dtype=t_.float64, device="cuda")
        body_final = t_.empty((nrel, K, 6), dtype=t_.float64, device="cuda")
        status = t_.empty(nrel, dtype=t_.int32, device="cuda")
        swd = dev.to_dev(sw)
        rel_d = dev.to_dev(np.ascontiguousarray(unq_t1s, dtype=np.float64))
        ns_d = dev.to_dev(ns, dtype=t_.int32)
        off_d = dev.to_dev(offsets, dtype=t_.int64)
        lib = _lib.load()
        handles = _lib.body_array([None if b is None else b.c_handle for b in bodies])
        _lib.check(lib.gamd_mockstream_dop853(pot.c_handle, handles, K, body_w.data_ptr(), nrel,
                                              rel_d.data_ptr(), ns_d.data_ptr(), off_d.data_ptr(),
                                              swd.data_ptr(), float(orbit_t[-1]),
                                              float(time[1] - time[0]), rtol, atol, nmax,
                                              out.data_ptr(), body_final.data_ptr(),
                                              status.data_ptr(), None, dev.stream_ptr()),
                   "gamd_mockstream_dop853")
        checked({"status": status}, "release systems")
        raw_stream = dev.to_host(out)[:ntot].T
        last = int(np.nonzero(ns)[0][-1])
        bf = dev.to_host(body_final)[last]  # (K, 6)
        stream = MockStream(pos=raw_stream[:3], vel=raw_stream[3:],
                            release_time=stream_w0.release_time,
                            lead_trail=stream_w0.lead_trail, frame=self.hamiltonian.frame)
        nbody_w = PhaseSpacePosition(pos=bf[:, :3].T, vel=bf[:, 3:].T,
                                     frame=self.hamiltonian.frame)
        return stream, nbody_w
