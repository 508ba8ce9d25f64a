"""MI355X engine for Mechanica's mdcore force-and-advance step.

The Python surface mirrors the parts of Mechanica's API that set up and run
the step (src/__init__.py, src/MxPotential.cpp:2459-2751,
src/MxUniverse.cpp:121-127):

    import mechanica_amd as m
    u = m.Universe(dim=[10, 10, 10], cells=[10, 10, 10], cutoff=1.0, dt=0.01)
    Argon = u.add_type(mass=39.4, radius=0.1)
    pot = m.Potential.lennard_jones_12_6(0.275, 1.0, 9.5075e-06, 6.1545e-03, 1e-3)
    u.bind(pot, Argon, Argon)
    u.load(positions, velocities, types)
    u.step(100)

Everything below calls the C ABI of libmxg.so (include/mxg.h); the step
runs only in the HIP kernels.
"""
import ctypes as C

import numpy as np

from . import _capi

NEWTONIAN = 0
OVERDAMPED = 1
# MxParticleFlags PARTICLE_FROZEN_X/Y/Z (MxParticle.h:58-60)
FROZEN_X = 1 << 3
FROZEN_Y = 1 << 4
FROZEN_Z = 1 << 5

ERR_TOOFAST = -29


class EngineError(RuntimeError):
    """Raised with the engine's error code and message (mdcore engine_err)."""

    def __init__(self, code, message):
        super().__init__("engine error %d: %s" % (code, message))
        self.code = code


def _dp(a):
    return a.ctypes.data_as(_capi.c_double_p)


def _fp(a):
    return a.ctypes.data_as(_capi.c_float_p)


def _ip(a):
    return a.ctypes.data_as(_capi.c_int_p)


class Potential:
    """A tabulated pair potential or a DPD potential (MxPotential)."""

    def __init__(self, handle, name):
        if not handle:
            raise ValueError("potential construction failed (potential_err %d)" % _capi.lib().mxg_potential_error())
        self._h = C.c_void_p(handle)
        self.name = name

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            try:
                _capi.lib().mxg_potential_free(h)
            except Exception:
                pass
            self._h = None

    # Factories with Mechanica's defaults (MxPotential.cpp:2459-2751).
    @staticmethod
    def lennard_jones_12_6(min, max, A, B, tol=None):
        if tol is None:
            tol = 0.001 * (max - min)
        return Potential(_capi.lib().mxg_potential_create_lj126(min, max, A, B, tol), "Lennard-Jones")

    @staticmethod
    def morse(d=1.0, a=6.0, r0=0.0, min=0.0001, max=3.0, tol=0.001):
        return Potential(_capi.lib().mxg_potential_create_morse(d, a, r0, min, max, tol), "Morse")

    @staticmethod
    def harmonic(k, r0, min=None, max=None, tol=None):
        if min is None:
            min = r0 - r0
        if max is None:
            max = r0 + r0
        if tol is None:
            tol = 0.01 * (max - min)
        return Potential(_capi.lib().mxg_potential_create_harmonic(min, max, k, r0, tol), "Harmonic")

    @staticmethod
    def harmonic_angle(k, theta0, min=0.0, max=np.pi, tol=1e-4):
        """potential_create_harmonic_angle: k (theta - theta0)^2 on angles in
        [min, max], tabulated in cos(theta) (MxPotential.cpp:555-597)."""
        return Potential(_capi.lib().mxg_potential_create_harmonic_angle(min, max, k, theta0, tol),
                         "Harmonic Angle")

    @staticmethod
    def harmonic_dihedral(k, n, delta, tol=1e-4):
        """potential_create_harmonic_dihedral: k (1 + cos(n phi - delta)),
        tabulated in cos(phi) (MxPotential.cpp:497-517)."""
        return Potential(_capi.lib().mxg_potential_create_harmonic_dihedral(k, int(n), delta, tol),
                         "Harmonic Dihedral")

    @staticmethod
    def dpd(alpha=1.0, gamma=1.0, sigma=1.0, cutoff=1.0, shifted=False):
        return Potential(_capi.lib().mxg_potential_create_dpd(alpha, gamma, sigma, cutoff, int(bool(shifted))),
                         "Dissapative Particle Dynamics")

    @property
    def info(self):
        inf = _capi.PotentialInfo()
        rc = _capi.lib().mxg_potential_get_info(self._h, C.byref(inf))
        if rc:
            raise ValueError("bad potential")
        return inf

    @property
    def kind(self):
        return self.info.kind

    @property
    def flags(self):
        return self.info.flags

    @property
    def n(self):
        return self.info.n

    @property
    def domain(self):
        inf = self.info
        return inf.a, inf.b

    @property
    def alpha(self):
        return np.array(self.info.alpha[:], dtype=np.float32)

    @property
    def r0_plusone(self):
        return self.info.r0_plusone

    @property
    def coefficients(self):
        L = _capi.lib()
        n = L.mxg_potential_coefficients(self._h, None, 0)
        out = np.zeros(max(n, 0), dtype=np.float32)
        if n > 0:
            L.mxg_potential_coefficients(self._h, _fp(out), n)
        return out

    def eval(self, r2):
        """potential_eval at r^2: (energy, force / r)."""
        e, f = C.c_float(), C.c_float()
        rc = _capi.lib().mxg_potential_eval(self._h, float(r2), C.byref(e), C.byref(f))
        if rc:
            raise ValueError("not a table potential")
        return e.value, f.value

    def eval_r(self, r):
        """potential_eval_r at r itself: (energy, dE/dr)."""
        e, f = C.c_float(), C.c_float()
        rc = _capi.lib().mxg_potential_eval_r(self._h, float(r), C.byref(e), C.byref(f))
        if rc:
            raise ValueError("not a table potential")
        return e.value, f.value

    def eval_ex(self, r2, ri=1.0, rj=1.0):
        """potential_eval_ex at r^2: (in_range, energy, force / r)."""
        e, f = C.c_float(), C.c_float()
        ok = _capi.lib().mxg_potential_eval_ex(self._h, float(ri), float(rj), float(r2), C.byref(e), C.byref(f))
        if ok < 0:
            raise ValueError("not a table potential")
        return bool(ok), e.value, f.value


FORWARD_EULER = 0
RUNGE_KUTTA_4 = 1

# BoundaryConditionKind (MxBoundaryConditions.hpp:15-23) and the sides.
BOUNDARY_VELOCITY = 1 << 0
BOUNDARY_PERIODIC = 1 << 1
BOUNDARY_FREESLIP = 1 << 2
BOUNDARY_POTENTIAL = 1 << 3
BOUNDARY_NO_SLIP = 1 << 4  # a velocity wall at rest
SIDES = ("left", "right", "front", "back", "bottom", "top")


class Universe:
    """One engine: a periodic universe (or a batch of them) on the GPU."""

    def __init__(self, dim=(10.0, 10.0, 10.0), cells=(4, 4, 4), cutoff=1.0, dt=0.01, origin=(0.0, 0.0, 0.0),
                 replicas=None):
        L = _capi.lib()
        h = C.c_void_p()
        if replicas is None:
            o = np.ascontiguousarray(origin, dtype=np.float64)
            d = np.ascontiguousarray(dim, dtype=np.float64)
            c = np.ascontiguousarray(cells, dtype=np.int32)
            rc = L.mxg_engine_init(C.byref(h), _dp(o), _dp(d), _ip(c), float(cutoff), float(dt))
            self.nreplicas = 1
        else:
            dims = np.ascontiguousarray([r[0] for r in replicas], dtype=np.float64).reshape(-1, 3)
            cl = np.ascontiguousarray([r[1] for r in replicas], dtype=np.int32).reshape(-1, 3)
            org = np.zeros_like(dims)
            rc = L.mxg_ensemble_init(C.byref(h), len(replicas), _dp(org), _dp(dims), _ip(cl), float(cutoff),
                                     float(dt))
            self.nreplicas = len(replicas)
        self._h = h
        self.cutoff = cutoff
        self.dt = dt
        self._pots = []
        self.N = 0
        self.M = 0
        self._check(rc)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.lib().mxg_engine_free(h)
            except Exception:
                pass
            self._h = None

    def _check(self, rc):
        if rc is not None and rc < 0:
            msg = _capi.lib().mxg_engine_error_message(self._h)
            raise EngineError(rc, msg.decode() if msg else "")
        return rc

    # set-up -------------------------------------------------------------
    def add_type(self, mass=1.0, radius=1.0, dynamics=NEWTONIAN):
        return self._check(_capi.lib().mxg_engine_addtype(self._h, float(mass), float(radius), int(dynamics)))

    def bind(self, pot, a, b):
        """m.bind(pot, A, B) -> engine_addpot (MxBind.cpp:18-60)."""
        self._pots.append(pot)
        self._check(_capi.lib().mxg_engine_addpot(self._h, pot._h, int(a), int(b)))

    def add_random_force(self, type_id, mean, std, duration):
        self._check(_capi.lib().mxg_engine_add_random_force(self._h, int(type_id), float(mean), float(std),
                                                            float(duration)))

    def add_berendsen(self, type_id, tau):
        """forces.berenderson_tstat(tau) bound to a type."""
        self._check(_capi.lib().mxg_engine_add_force_berendsen(self._h, int(type_id), float(tau)))

    def add_constant_force(self, type_id, force):
        f = np.ascontiguousarray(force, dtype=np.float32).reshape(3)
        self._check(_capi.lib().mxg_engine_add_force_constant(self._h, int(type_id), _fp(f)))

    def add_friction(self, type_id, coef, mean=0.0, std=0.0, duration=0.01):
        """forces.friction(coef, mean, std, durration) bound to a type."""
        self._check(_capi.lib().mxg_engine_add_force_friction(self._h, int(type_id), float(coef), float(mean),
                                                              float(std), float(duration)))

    def set_target_energy(self, type_id, energy):
        """MxParticleType.target_energy (the Berendsen target)."""
        self._check(_capi.lib().mxg_engine_set_target_energy(self._h, int(type_id), float(energy)))

    def set_seed(self, seed):
        """Seed of the counter-based generators; resets every replica's seed
        to seed + r * 0x9E3779B97F4A7C15."""
        self._check(_capi.lib().mxg_engine_set_seed(self._h, int(seed) & 0xFFFFFFFFFFFFFFFF))

    def set_replica_seed(self, replica, seed):
        """Seed of one replica's random streams (a replica evolves as a
        standalone engine with the same particles and this seed)."""
        self._check(_capi.lib().mxg_engine_set_replica_seed(self._h, int(replica),
                                                            int(seed) & 0xFFFFFFFFFFFFFFFF))

    def replica_seed(self, replica):
        s = C.c_uint64()
        self._check(_capi.lib().mxg_engine_replica_seed(self._h, int(replica), C.byref(s)))
        return s.value

    def set_dpd_noise(self, mode):
        self._check(_capi.lib().mxg_engine_set_dpd_noise(self._h, int(mode)))

    def set_max_dist_fraction(self, fraction):
        self._check(_capi.lib().mxg_engine_set_max_dist_fraction(self._h, float(fraction)))

    def set_brick(self, bx=0, by=0, bz=0, cap=0, pmax=0):
        """Brick shape of the force kernel (0 = automatic)."""
        self._check(_capi.lib().mxg_engine_set_brick(self._h, int(bx), int(by), int(bz), int(cap), int(pmax)))

    def brick(self):
        d = np.zeros(3, np.int32)
        cap, pmax = C.c_int(), C.c_int()
        self._check(_capi.lib().mxg_engine_get_brick(self._h, _ip(d), C.byref(cap), C.byref(pmax)))
        return tuple(int(x) for x in d), cap.value, pmax.value

    def autotune(self, steps_per=5):
        """Picks the fastest brick layout using real steps (the trajectory
        is the same for every layout).  Returns the steps taken, the chosen
        layout and the timings."""
        L = _capi.lib()
        taken = C.c_int()
        self._check(L.mxg_engine_autotune(self._h, int(steps_per), C.byref(taken)))
        n = L.mxg_engine_tune_log(self._h, 0, None, None)
        lay = np.zeros((max(n, 1), 5), np.int32)
        ms = np.zeros(max(n, 1), np.float64)
        L.mxg_engine_tune_log(self._h, n, _ip(lay), _dp(ms))
        log = [(tuple(int(v) for v in lay[k]), float(ms[k])) for k in range(n)]
        return {"steps": taken.value, "brick": self.brick(), "waves": self.waves(), "log": log}

    def set_waves(self, waves):
        """Waves per SIMD of the force kernel: 3 or 4 (4 applies to steps
        with bonds or terms)."""
        self._check(_capi.lib().mxg_engine_set_waves(self._h, int(waves)))

    def waves(self):
        return int(_capi.lib().mxg_engine_get_waves(self._h))

    def set_integrator(self, kind):
        """FORWARD_EULER (0) or RUNGE_KUTTA_4 (1), as Mechanica's integrator."""
        self._check(_capi.lib().mxg_engine_set_integrator(self._h, int(kind)))

    def set_graphs(self, on=True):
        """Replay pairs of steps from captured HIP graphs (default on)."""
        self._check(_capi.lib().mxg_engine_set_graphs(self._h, int(bool(on))))

    def set_virial(self, on=True):
        self._check(_capi.lib().mxg_engine_set_virial(self._h, int(bool(on))))

    def load(self, x, v=None, types=None, replica=None, flags=None, pid=None):
        """engine_load: particle k gets id pid[k] (k without pid); bonds and
        terms name particles by id, read-backs are in load order."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        N = x.shape[0]
        v = np.zeros_like(x) if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        t = np.zeros(N, np.int32) if types is None else np.ascontiguousarray(types, dtype=np.int32)
        r = None if replica is None else np.ascontiguousarray(replica, dtype=np.int32)
        p = None if pid is None else np.ascontiguousarray(pid, dtype=np.int32)
        fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint16)
        self._check(_capi.lib().mxg_engine_load(
            self._h, N, _dp(x), _dp(v), _ip(t), _ip(p) if p is not None else None,
            _ip(r) if r is not None else None,
            fl.ctypes.data_as(_capi.c_u16_p) if fl is not None else None))
        self.N = N
        self.M = 0

    def set_bonds(self, i, j, pots, dissociation=None):
        i = np.ascontiguousarray(i, dtype=np.int32)
        j = np.ascontiguousarray(j, dtype=np.int32)
        M = len(i)
        if isinstance(pots, Potential):
            pots = [pots] * M
        self._pots.extend(set(pots))
        arr = (C.c_void_p * max(M, 1))(*[p._h.value for p in pots])
        d = None
        if dissociation is not None:
            d = np.ascontiguousarray(np.broadcast_to(np.asarray(dissociation, np.float64), (M,)))
        self._check(_capi.lib().mxg_engine_set_bonds(self._h, M, _ip(i), _ip(j), arr,
                                                     _dp(d) if d is not None else None))
        self.M = M

    def set_boundaries(self, kinds, velocity=None, restore=None, radius=None):
        """MxBoundaryConditions: kinds is a dict or sequence over the sides
        left, right, front, back, bottom, top of BOUNDARY_* bits or the
        names 'periodic', 'free_slip', 'no_slip', 'potential' (potential
        walls also reflect, as the reference's 'potential' string does);
        velocity, restore and radius are per-side dicts or sequences."""
        names = {"periodic": BOUNDARY_PERIODIC, "free_slip": BOUNDARY_FREESLIP,
                 "freeslip": BOUNDARY_FREESLIP, "no_slip": BOUNDARY_VELOCITY, "noslip": BOUNDARY_VELOCITY,
                 "velocity": BOUNDARY_VELOCITY, "potential": BOUNDARY_POTENTIAL | BOUNDARY_FREESLIP}

        def per_side(v, default):
            if v is None:
                return [default] * 6
            if isinstance(v, dict):
                return [v.get(sd, default) for sd in SIDES]
            return list(v)

        k = [names[x.lower()] if isinstance(x, str) else int(x) for x in per_side(kinds, BOUNDARY_PERIODIC)]
        vel = np.ascontiguousarray(np.asarray(per_side(velocity, (0.0, 0.0, 0.0)), np.float32).reshape(6, 3))
        res = np.ascontiguousarray(per_side(restore, 1.0), np.float32)
        rad = np.ascontiguousarray(per_side(radius, 0.0), np.float32)
        karr = np.ascontiguousarray(k, np.int32)
        self._check(_capi.lib().mxg_engine_set_boundaries(self._h, _ip(karr), _fp(vel), _fp(res), _fp(rad)))

    def set_boundary_potential(self, side, type_id, pot):
        """MxBoundaryCondition.set_potential for one side ('left'... or 0-5)."""
        sd = SIDES.index(side) if isinstance(side, str) else int(side)
        if pot is not None:
            self._pots.append(pot)
        self._check(_capi.lib().mxg_engine_set_boundary_potential(
            self._h, sd, int(type_id), pot._h.value if pot is not None else None))

    # species and fluxes ----------------------------------------------------
    FICK, SECRETE, UPTAKE = 0, 1, 2

    def set_species(self, type_id, constant_mask=0):
        """Up to four species per particle of the type; bit k of
        constant_mask keeps species k constant."""
        self._check(_capi.lib().mxg_engine_set_species(self._h, int(type_id), int(constant_mask)))

    def set_species_values(self, values):
        """Species values by particle id (N x up to 4)."""
        a = np.asarray(values, np.float32)
        v = np.zeros((self.N, 4), np.float32)
        v[:, :a.shape[1]] = a
        self._check(_capi.lib().mxg_engine_set_species_values(self._h, _fp(v)))

    def species(self):
        """(values, pending flux accumulators), N x 4 each, by id."""
        v = np.zeros((self.N, 4), np.float32)
        q = np.zeros((self.N, 4), np.float32)
        self._check(_capi.lib().mxg_engine_get_species(self._h, _fp(v), _fp(q)))
        return v, q

    def add_flux(self, kind, type_a, type_b, species_a, species_b, k, decay=0.0, target=0.0):
        """MxFluxes_FluxEx: kind FICK, SECRETE or UPTAKE."""
        self._check(_capi.lib().mxg_engine_add_flux(self._h, int(kind), int(type_a), int(type_b), int(species_a),
                                                    int(species_b), float(k), float(decay), float(target)))

    def _term_pots(self, pots, M):
        if isinstance(pots, Potential):
            pots = [pots] * M
        pots = list(pots)
        if len(pots) != M:
            raise ValueError("one potential per term")
        self._pots.extend(set(pots))
        return (C.c_void_p * max(M, 1))(*[p._h.value for p in pots])

    def set_angles(self, i, j, k, pots):
        """engine_angle_add for each (i[a], j[a], k[a]), j the vertex;
        replaces the previous angles."""
        ids = [np.ascontiguousarray(a, dtype=np.int32) for a in (i, j, k)]
        M = len(ids[0])
        arr = self._term_pots(pots, M)
        self._check(_capi.lib().mxg_engine_set_angles(self._h, M, *[_ip(a) for a in ids], arr))

    def set_dihedrals(self, i, j, k, l, pots):
        """engine_dihedral_add for each (i, j, k, l) about the axis j-k;
        replaces the previous dihedrals."""
        ids = [np.ascontiguousarray(a, dtype=np.int32) for a in (i, j, k, l)]
        M = len(ids[0])
        arr = self._term_pots(pots, M)
        self._check(_capi.lib().mxg_engine_set_dihedrals(self._h, M, *[_ip(a) for a in ids], arr))

    def set_exclusions(self, i, j):
        """engine_exclusion_add for each pair; replaces the previous set."""
        i = np.ascontiguousarray(i, dtype=np.int32)
        j = np.ascontiguousarray(j, dtype=np.int32)
        self._check(_capi.lib().mxg_engine_set_exclusions(self._h, len(i), _ip(i), _ip(j)))

    def term_energies(self, replica=-1):
        """Energies of the angles, dihedrals and exclusions of the last force
        evaluation."""
        out = np.zeros(3, np.float64)
        self._check(_capi.lib().mxg_engine_term_energies(self._h, int(replica), _dp(out)))
        return {"epot_angle": out[0], "epot_dihedral": out[1], "epot_exclusion": out[2]}

    # stepping -----------------------------------------------------------
    def step(self, nsteps=1):
        """engine_step, nsteps times (MxUniverse_Step)."""
        return self._check(_capi.lib().mxg_engine_step(self._h, int(nsteps)))

    def sync(self):
        self._check(_capi.lib().mxg_engine_sync(self._h))

    # read-back ----------------------------------------------------------
    def state(self):
        """Per-particle arrays by id: local x, cell, v, f, persistent force."""
        N = self.N
        xl = np.zeros((N, 3), np.float32)
        cell = np.zeros(N, np.int32)
        v = np.zeros((N, 3), np.float32)
        f = np.zeros((N, 3), np.float32)
        pf = np.zeros((N, 3), np.float32)
        self._check(_capi.lib().mxg_engine_state(self._h, _fp(xl), _ip(cell), _fp(v), _fp(f), _fp(pf)))
        return {"x": xl, "cell": cell, "v": v, "f": f, "pf": pf}

    def unload(self):
        x = np.zeros((self.N, 3), np.float64)
        v = np.zeros((self.N, 3), np.float64)
        self._check(_capi.lib().mxg_engine_unload(self._h, _dp(x), _dp(v)))
        return x, v

    def number_density(self):
        nd = np.zeros(self.N, np.float32)
        self._check(_capi.lib().mxg_engine_number_density(self._h, _fp(nd)))
        return nd

    def neighbour_counts(self):
        c = np.zeros(self.N, np.int32)
        self._check(_capi.lib().mxg_engine_neighbour_counts(self._h, _ip(c)))
        return c

    def observables(self, replica=-1):
        out = np.zeros(12, np.float64)
        self._check(_capi.lib().mxg_engine_observables(self._h, int(replica), _dp(out)))
        return {"epot_nonbond": out[0], "epot_bond": out[1], "ekin": out[2], "virial": out[3:].reshape(3, 3)}

    def bond_active(self):
        a = np.zeros(max(self.M, 1), np.uint8)
        self._check(_capi.lib().mxg_engine_bond_active(self._h, a.ctypes.data_as(_capi.c_u8_p)))
        return a[: self.M].astype(bool)

    def geometry(self, replica=0):
        cd = np.zeros(3, np.int32)
        h = np.zeros(3, np.float64)
        self._check(_capi.lib().mxg_engine_geometry(self._h, int(replica), _ip(cd), _dp(h)))
        return cd, h

    @property
    def time(self):
        t = C.c_uint64()
        _capi.lib().mxg_engine_time(self._h, C.byref(t))
        return t.value

    @property
    def error(self):
        return _capi.lib().mxg_engine_error(self._h)

    def set_timing(self, on=True):
        _capi.lib().mxg_engine_set_timing(self._h, int(bool(on)))

    def timing(self):
        a, b = C.c_double(), C.c_double()
        _capi.lib().mxg_engine_timing(self._h, C.byref(a), C.byref(b))
        return a.value, b.value


def device_count():
    return _capi.lib().mxg_device_count()
