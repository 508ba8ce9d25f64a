"""ORACLE — test infrastructure only.

ctypes wrapper of liboracle.so, the CPU restatement of mdcore's step
(mdcore_oracle.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, as the checker or as the timed CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
_fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
_ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))

_SIG = {
    "oracle_create": (C.c_void_p, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_double,
                                   C.c_double, C.c_int]),
    "oracle_free": (None, [C.c_void_p]),
    "oracle_geometry": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "oracle_set_threads": (None, [C.c_void_p, C.c_int]),
    "oracle_set_seed": (None, [C.c_void_p, C.c_uint64]),
    "oracle_set_max_dist_fraction": (None, [C.c_void_p, C.c_float]),
    "oracle_philox": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "oracle_set_dpd_counter_noise": (None, [C.c_void_p, C.c_int]),
    "oracle_set_time": (None, [C.c_void_p, C.c_uint64]),
    "oracle_add_type": (C.c_int, [C.c_void_p, C.c_double, C.c_float, C.c_int]),
    "oracle_set_random_force": (C.c_int, [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float]),
    "oracle_add_pot": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.POINTER(C.c_float),
                                 C.c_float, C.c_float, C.c_float, C.c_int, C.POINTER(C.c_float), C.c_float,
                                 C.c_float, C.c_float]),
    "oracle_add_bond_pot": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_float, C.c_float, C.c_int,
                                      C.POINTER(C.c_float)]),
    "oracle_load": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(C.c_int), C.POINTER(C.c_uint16)]),
    "oracle_load_local": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int),
                                    C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_uint16),
                                    C.POINTER(C.c_float)]),
    "oracle_set_bonds": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "oracle_set_boundaries": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_float),
                                        C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "oracle_set_boundary_pot": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_uint32,
                                          C.POINTER(C.c_float), C.c_float, C.c_float, C.c_float, C.c_int,
                                          C.POINTER(C.c_float), C.c_float, C.c_float, C.c_float]),
    "oracle_add_angle_pot": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_float, C.c_float, C.c_int,
                                       C.POINTER(C.c_float)]),
    "oracle_add_dihedral_pot": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_float, C.c_float, C.c_int,
                                          C.POINTER(C.c_float)]),
    "oracle_set_angles": (C.c_int, [C.c_void_p, C.c_int] + [C.POINTER(C.c_int)] * 4),
    "oracle_set_dihedrals": (C.c_int, [C.c_void_p, C.c_int] + [C.POINTER(C.c_int)] * 5),
    "oracle_set_exclusions": (C.c_int, [C.c_void_p, C.c_int] + [C.POINTER(C.c_int)] * 2),
    "oracle_term_energies": (None, [C.c_void_p, C.POINTER(C.c_double)]),
    "oracle_step": (C.c_int, [C.c_void_p, C.c_int]),
    "oracle_set_integrator": (None, [C.c_void_p, C.c_int]),
    "oracle_set_single_body": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "oracle_set_target_energy": (None, [C.c_void_p, C.c_int, C.c_double]),
    "oracle_force": (C.c_int, [C.c_void_p]),
    "oracle_step_with_forces": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "oracle_get_state": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_float),
                                   C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    "oracle_cell_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "oracle_global_positions": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "oracle_observables": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "oracle_bond_active": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint8)]),
    "oracle_set_species": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "oracle_add_flux": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                  C.c_float]),
    "oracle_set_species_values": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "oracle_get_species": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "oracle_virial64": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "oracle_forces64": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "oracle_error": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "oracle_number_density": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
}

_lib = None


def build():
    """Compiles liboracle.so with the recipe in oracle/Makefile."""
    r = subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for k, (res, args) in _SIG.items():
            f = getattr(L, k)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _pot_tables(pot):
    inf = pot.info
    alpha = np.array(inf.alpha[:], np.float32)
    c = pot.coefficients
    return inf, alpha, c


class Oracle:
    def __init__(self, dim, cells, cutoff, dt, threads=1, origin=(0.0, 0.0, 0.0)):
        L = lib()
        o = np.ascontiguousarray(origin, np.float64)
        d = np.ascontiguousarray(dim, np.float64)
        c = np.ascontiguousarray(cells, np.int32)
        self._h = L.oracle_create(_dp(o), _dp(d), _ip(c), float(cutoff), float(dt), int(threads))
        if not self._h:
            raise ValueError("oracle geometry rejected")
        self.N = 0
        self.M = 0

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_free(self._h)
            self._h = None

    def geometry(self):
        cd = np.zeros(3, np.int32)
        h = np.zeros(3, np.float64)
        nt = C.c_int()
        lib().oracle_geometry(self._h, _ip(cd), _dp(h), C.byref(nt))
        return cd, h, nt.value

    def set_threads(self, n):
        lib().oracle_set_threads(self._h, int(n))

    def set_seed(self, s):
        lib().oracle_set_seed(self._h, int(s) & 0xFFFFFFFFFFFFFFFF)

    def set_max_dist_fraction(self, f):
        lib().oracle_set_max_dist_fraction(self._h, float(f))

    def set_dpd_noise(self, on):
        lib().oracle_set_dpd_counter_noise(self._h, int(on))

    def add_type(self, mass, radius, dynamics):
        return lib().oracle_add_type(self._h, float(mass), float(radius), int(dynamics))

    def set_single_body(self, t, kind, params):
        p = np.ascontiguousarray(params, np.float32)
        lib().oracle_set_single_body(self._h, int(t), int(kind), p.ctypes.data_as(C.POINTER(C.c_float)))

    def set_target_energy(self, t, e):
        lib().oracle_set_target_energy(self._h, int(t), float(e))

    def add_random_force(self, t, mean, std, duration):
        lib().oracle_set_random_force(self._h, int(t), float(mean), float(std), float(duration))

    def bind(self, pot, ti, tj):
        inf, alpha, c = _pot_tables(pot)
        cp = _fp(c) if c.size else None
        lib().oracle_add_pot(self._h, int(ti), int(tj), inf.kind, inf.flags, _fp(alpha), inf.a, inf.b,
                             inf.r0_plusone, inf.n, cp, inf.dpd_alpha, inf.dpd_gamma, inf.dpd_sigma)

    def set_boundaries(self, kinds, velocity, restore, radius):
        """Six sides: kind bits, velocities (6 x 3), restore and radius."""
        k = np.ascontiguousarray(kinds, np.int32)
        vel = np.ascontiguousarray(velocity, np.float32).reshape(-1)
        res = np.ascontiguousarray(restore, np.float32)
        rad = np.ascontiguousarray(radius, np.float32)
        rc = lib().oracle_set_boundaries(self._h, _ip(k), _fp(vel), _fp(res), _fp(rad))
        if rc:
            raise ValueError("oracle boundaries failed %d" % rc)

    def set_boundary_potential(self, side, ti, pot):
        inf, alpha, c = _pot_tables(pot)
        cp = _fp(c) if c.size else None
        lib().oracle_set_boundary_pot(self._h, int(side), int(ti), inf.kind, inf.flags, _fp(alpha), inf.a, inf.b,
                                      inf.r0_plusone, inf.n, cp, inf.dpd_alpha, inf.dpd_gamma, inf.dpd_sigma)

    def load(self, x, v, types, flags=None):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(types, np.int32)
        fl = None if flags is None else np.ascontiguousarray(flags, np.uint16)
        rc = lib().oracle_load(self._h, x.shape[0], _dp(x), _dp(v), _ip(t),
                               fl.ctypes.data_as(C.POINTER(C.c_uint16)) if fl is not None else None)
        if rc:
            raise ValueError("oracle load failed %d" % rc)
        self.N = x.shape[0]

    def load_local(self, xl, cell, v, types, flags=None, pf=None):
        """Loads a cell-local state (as the engine reports it) without going
        through global coordinates: particle i in cell[i] at xl[i]."""
        xl = np.ascontiguousarray(xl, np.float32).reshape(-1, 3)
        c = np.ascontiguousarray(cell, np.int32)
        v = np.ascontiguousarray(v, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(types, np.int32)
        fl = None if flags is None else np.ascontiguousarray(flags, np.uint16)
        p = None if pf is None else np.ascontiguousarray(pf, np.float32).reshape(-1, 3)
        rc = lib().oracle_load_local(self._h, xl.shape[0], _fp(xl), _ip(c), _fp(v), _ip(t),
                                     fl.ctypes.data_as(C.POINTER(C.c_uint16)) if fl is not None else None,
                                     _fp(p) if p is not None else None)
        if rc:
            raise ValueError("oracle local load failed %d" % rc)
        self.N = xl.shape[0]

    def set_bonds(self, i, j, pot, dissociation):
        inf, alpha, c = _pot_tables(pot)
        pidx = lib().oracle_add_bond_pot(self._h, _fp(alpha), inf.a, inf.b, inf.n, _fp(c))
        i = np.ascontiguousarray(i, np.int32)
        j = np.ascontiguousarray(j, np.int32)
        M = len(i)
        p = np.full(M, pidx, np.int32)
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dissociation, np.float64), (M,)))
        lib().oracle_set_bonds(self._h, M, _ip(i), _ip(j), _ip(p), _dp(d))
        self.M = M

    def set_angles(self, i, j, k, pot):
        inf, alpha, c = _pot_tables(pot)
        pidx = lib().oracle_add_angle_pot(self._h, _fp(alpha), inf.a, inf.b, inf.n, _fp(c))
        ids = [np.ascontiguousarray(a, np.int32) for a in (i, j, k)]
        p = np.full(len(ids[0]), pidx, np.int32)
        lib().oracle_set_angles(self._h, len(p), *[_ip(a) for a in ids], _ip(p))

    def set_dihedrals(self, i, j, k, l, pot):
        inf, alpha, c = _pot_tables(pot)
        pidx = lib().oracle_add_dihedral_pot(self._h, _fp(alpha), inf.a, inf.b, inf.n, _fp(c))
        ids = [np.ascontiguousarray(a, np.int32) for a in (i, j, k, l)]
        p = np.full(len(ids[0]), pidx, np.int32)
        lib().oracle_set_dihedrals(self._h, len(p), *[_ip(a) for a in ids], _ip(p))

    def set_exclusions(self, i, j):
        i = np.ascontiguousarray(i, np.int32)
        j = np.ascontiguousarray(j, np.int32)
        lib().oracle_set_exclusions(self._h, len(i), _ip(i), _ip(j))

    def term_energies(self):
        out = np.zeros(3, np.float64)
        lib().oracle_term_energies(self._h, _dp(out))
        return {"epot_angle": out[0], "epot_dihedral": out[1], "epot_exclusion": out[2]}

    def set_integrator(self, kind):
        """0: forward Euler, 1: Runge-Kutta 4 (engine_advance_runge_kutta_4)."""
        lib().oracle_set_integrator(self._h, int(kind))

    def step(self, n=1):
        return lib().oracle_step(self._h, int(n))

    def force(self):
        return lib().oracle_force(self._h)

    def step_with_forces(self, f):
        f = np.ascontiguousarray(f, np.float32).reshape(-1)
        return lib().oracle_step_with_forces(self._h, _fp(f))

    def state(self):
        N = self.N
        xl = np.zeros((N, 3), np.float32)
        cell = np.zeros(N, np.int32)
        v = np.zeros((N, 3), np.float32)
        f = np.zeros((N, 3), np.float32)
        nc = np.zeros(N, np.int32)
        pf = np.zeros((N, 3), np.float32)
        lib().oracle_get_state(self._h, _fp(xl), _ip(cell), _fp(v), _fp(f), _ip(nc), _fp(pf))
        return {"x": xl, "cell": cell, "v": v, "f": f, "ncount": nc, "pf": pf}

    def number_density(self):
        nd = np.zeros(self.N, np.float32)
        lib().oracle_number_density(self._h, _fp(nd))
        return nd

    def global_positions(self):
        x = np.zeros((self.N, 3), np.float64)
        lib().oracle_global_positions(self._h, _dp(x))
        return x

    def observables(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        lib().oracle_observables(self._h, C.byref(a), C.byref(b), C.byref(c))
        return {"epot_nonbond": a.value, "epot_bond": b.value, "ekin": c.value}

    def bond_active(self):
        a = np.zeros(max(self.M, 1), np.uint8)
        lib().oracle_bond_active(self._h, a.ctypes.data_as(C.POINTER(C.c_uint8)))
        return a[: self.M].astype(bool)

    def forces64(self):
        F = np.zeros((self.N, 3), np.float64)
        e = C.c_double()
        lib().oracle_forces64(self._h, _dp(F), C.byref(e))
        return F, e.value

    def set_species(self, t, constant_mask=0):
        lib().oracle_set_species(self._h, int(t), int(constant_mask))

    def add_flux(self, kind, ta, tb, ia, ib, k, decay=0.0, target=0.0):
        lib().oracle_add_flux(self._h, int(kind), int(ta), int(tb), int(ia), int(ib), float(k), float(decay),
                              float(target))

    def set_species_values(self, values):
        v = np.zeros((self.N, 4), np.float32)
        a = np.asarray(values, np.float32)
        v[:, :a.shape[1]] = a
        lib().oracle_set_species_values(self._h, _fp(v))

    def species(self):
        v = np.zeros((self.N, 4), np.float32)
        q = np.zeros((self.N, 4), np.float32)
        lib().oracle_get_species(self._h, _fp(v), _fp(q))
        return v, q

    def virial64(self):
        """FP64 virial tensor sum (x_i - x_j) (x) f_i over pairs and bonds."""
        W = np.zeros(9, np.float64)
        lib().oracle_virial64(self._h, _dp(W))
        return W.reshape(3, 3)

    def error(self):
        pid = C.c_int()
        rc = lib().oracle_error(self._h, C.byref(pid))
        return rc, pid.value


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    out = (C.c_uint32 * 4)()
    lib().oracle_philox(c, k, out)
    return list(out)


def from_workload(w, m, threads=1, seed=None):
    """Builds an Oracle for workload dict w, with tables from module m."""
    from mechanica_amd import workloads as W
    o = Oracle(w["dim"], w["cells"], w["cutoff"], w["dt"], threads=threads)
    for t in w["types"]:
        o.add_type(t["mass"], t["radius"], t["dynamics"])
    keep = []
    for spec in w["pots"]:
        p = W.make_potential(m, spec)
        o.bind(p, spec[2], spec[3])
        keep.append(p)
    if w.get("boundaries"):
        b = W.boundary_arrays(w["boundaries"])
        o.set_boundaries(*b)
    for side, t, spec in w.get("wall_pots") or ():
        p = W.make_potential(m, spec)
        o.set_boundary_potential(side, t, p)
        keep.append(p)
    if seed is not None:
        o.set_seed(seed)
    if w.get("random"):
        rf = w["random"]
        for t in range(len(w["types"])):
            o.add_random_force(t, rf["mean"], rf["std"], rf["duration"])
    if w.get("dpd_noise"):
        o.set_dpd_noise(1)
    if w.get("integrator"):
        o.set_integrator(w["integrator"])
    for t, (kind, params) in (w.get("single_body") or {}).items():
        o.set_single_body(t, kind, params)
    for t, e in (w.get("target_energy") or {}).items():
        o.set_target_energy(t, e)
    o.load(w["x"], w["v"], w["type"], w.get("flags"))
    if w.get("bonds"):
        b = w["bonds"]
        bp = W.make_potential(m, b["pot"])
        o.set_bonds(b["i"], b["j"], bp, b["dissociation"])
        keep.append(bp)
    for t, mask in (w.get("species_const") or {}).items():
        o.set_species(t, mask)
    for fx in w.get("fluxes") or ():
        o.add_flux(*fx)
    if w.get("species") is not None:
        o.set_species_values(w["species"])
    if w.get("exclusions"):
        o.set_exclusions(w["exclusions"]["i"], w["exclusions"]["j"])
    if w.get("angles"):
        a = w["angles"]
        ap = W.make_potential(m, a["pot"])
        o.set_angles(a["i"], a["j"], a["k"], ap)
        keep.append(ap)
    if w.get("dihedrals"):
        d = w["dihedrals"]
        dp = W.make_potential(m, d["pot"])
        o.set_dihedrals(d["i"], d["j"], d["k"], d["l"], dp)
        keep.append(dp)
    o._keep = keep
    return o
