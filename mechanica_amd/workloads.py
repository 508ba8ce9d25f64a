"""Seeded synthetic inputs for the benchmark workloads.

Each builder returns a dict with the universe geometry, particle arrays,
types, potentials (as constructor arguments, so the same tables can be built
for the GPU engine and handed to the CPU checker) and bonds.  Sizes are
parameters so tests can build the same workload at small N.

  argon_lj   N LJ 12-6 argon particles (mass 39.4, radius 0.1) uniform in a
             periodic box at density 10, Gaussian velocities at kT = 1,
             LJ(0.275, 1.0, A=9.5075e-6, B=6.1545e-3, tol=1e-3), cutoff 1,
             Euler dt = 0.01 (examples/argon.py).
  dpd_fluid  N DPD particles at number density 3 (box side (N/3)^(1/3)),
             alpha = 25, gamma = 4.5, sigma = 0 or 3, cutoff 1, dt = 0.01,
             cells of side ~1 (examples/dpd1.py, dpd2.py at scale).
  lattice    n^3 simple-cubic lattice (spacing 1, jitter U(+-0.05)) with
             harmonic bonds to +x/+y/+z neighbours inside the box (the open
             lattice of src/lattice.py:449-522; 3 n^2 (n-1) bonds), K = 100,
             r0 = 1, dissociation energy 5, Morse(d=1, a=6, r0=1, max=2)
             between all pairs, cutoff 2, overdamped, dt = 0.005.
  molecules  n^3 four-bead chains (bond 0.3, angles near 110 deg, random
             dihedrals) centred on a cubic lattice of spacing 1.6 (jitter
             U(+-0.15)), with harmonic bonds 1-2, 2-3, 3-4 (K 100, r0 0.3),
             harmonic angles 1-2-3, 2-3-4 (K 20, theta0 110 deg, defined on
             theta0 +- 60 deg), a harmonic dihedral 1-2-3-4 (K 1, n 3,
             delta 0 at tolerance 1e-4 or pi/4 at 1e-2), exclusions of the 1-2 and 1-3 pairs, and LJ
             12-6 (sigma 0.25, eps 0.5, table on [0.1, 1]) between all
             particles; cutoff 1,
             Newtonian, mass 1, dt 0.002 (the bonded set of the reference's
             flexible-water example, examples/flexible/flexible.cpp:304-336,
             on a non-Coulomb pair potential).
  channel    examples/dpd1.py at scale: DPD (alpha 1, gamma 1, sigma 1.5,
             cutoff 1) particles of mass 10 at number density 1.875 in a
             15:12:10 box, no-slip walls in z, periodic in x and y, driven
             by a constant force (0.01, 0, 0), dt 0.1.
  species    the DPD fluid at density 3 in two types carrying three
             species, with Fick fluxes within and between the types,
             secretion and uptake (Flux.cpp, flux_eval.hpp).
  cell_sort  two-type Morse cell-sorting universe (examples/cell_sorting.py)
             at density 0.8: AA and BB d = 3, AB d = 1, a = 6, r0 = 1,
             cutoff 2.5, overdamped with a persistent random force whose
             per-component variance is 2 kT / (dt / m) at kT = 0.1.
"""
import math

import numpy as np

LJ_A = 9.5075e-06
LJ_B = 6.1545e-03


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _min_dist_fill(rng, n, L, dmin, max_tries=200):
    """Uniform positions in [0, L)^3 with all periodic pair distances >= dmin
    (rejection on a cell grid)."""
    pts = np.zeros((n, 3))
    nc = max(1, int(L / dmin))
    h = L / nc
    grid = {}
    k = 0
    tries = 0
    while k < n:
        p = rng.random(3) * L
        ci = tuple((p / h).astype(int) % nc)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    cc = ((ci[0] + dx) % nc, (ci[1] + dy) % nc, (ci[2] + dz) % nc)
                    for q in grid.get(cc, ()):
                        d = pts[q] - p
                        d -= L * np.round(d / L)
                        if d @ d < dmin * dmin:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            pts[k] = p
            grid.setdefault(ci, []).append(k)
            k += 1
            tries = 0
        else:
            tries += 1
            if tries > max_tries * n:
                raise RuntimeError("could not place particles")
    return pts


def argon_lj(N=10000, density=10.0, seed=1, min_dist=None):
    rng = _rng(seed)
    L = (N / density) ** (1.0 / 3.0)
    cells = max(3, int(math.floor(L / 1.0)))
    mass = 39.4
    if min_dist:
        x = _min_dist_fill(rng, N, L, min_dist)
    else:
        x = rng.random((N, 3)) * L
    v = rng.normal(0.0, math.sqrt(1.0 / mass), size=(N, 3))
    return {
        "name": "argon_lj",
        "dim": [L, L, L], "cells": [cells] * 3, "cutoff": 1.0, "dt": 0.01,
        "types": [{"mass": mass, "radius": 0.1, "dynamics": 0}],
        "x": x, "v": v, "type": np.zeros(N, np.int32),
        "pots": [("lj126", (0.275, 1.0, LJ_A, LJ_B, 1e-3), 0, 0)],
        "bonds": None,
    }


def dpd_fluid(N=1048576, density=3.0, sigma=0.0, seed=2):
    rng = _rng(seed)
    L = (N / density) ** (1.0 / 3.0)
    cells = max(3, int(math.floor(L / 1.0)))
    x = rng.random((N, 3)) * L
    v = rng.normal(0.0, 1.0, size=(N, 3))
    return {
        "name": "dpd_fluid",
        "dim": [L, L, L], "cells": [cells] * 3, "cutoff": 1.0, "dt": 0.01,
        "types": [{"mass": 1.0, "radius": 1.0, "dynamics": 0}],
        "x": x, "v": v, "type": np.zeros(N, np.int32),
        "pots": [("dpd", (25.0, 4.5, sigma, 1.0, 0), 0, 0)],
        "bonds": None,
    }


def lattice(n=64, seed=3, jitter=0.05, dissociation=5.0):
    rng = _rng(seed)
    idx = np.arange(n)
    g = np.stack(np.meshgrid(idx, idx, idx, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    x = g + 0.5 + rng.uniform(-jitter, jitter, size=g.shape)
    N = n ** 3
    lin = np.arange(N).reshape(n, n, n)
    bi, bj = [], []
    for ax in range(3):
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[ax] = slice(0, n - 1)
        sl_b[ax] = slice(1, n)
        bi.append(lin[tuple(sl_a)].ravel())
        bj.append(lin[tuple(sl_b)].ravel())
    bi = np.concatenate(bi).astype(np.int32)
    bj = np.concatenate(bj).astype(np.int32)
    cells = max(3, n // 2)
    return {
        "name": "lattice",
        "dim": [float(n)] * 3, "cells": [cells] * 3, "cutoff": 2.0, "dt": 0.005,
        "types": [{"mass": 1.0, "radius": 1.0, "dynamics": 1}],
        "x": x, "v": np.zeros_like(x), "type": np.zeros(N, np.int32),
        "pots": [("morse", (1.0, 6.0, 1.0, 1e-4, 2.0, 1e-3), 0, 0)],
        "bonds": {"i": bi, "j": bj, "pot": ("harmonic", (0.0, 2.0, 100.0, 1.0, 0.02)),
                  "dissociation": dissociation},
    }


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _chains(rng, M, bond=0.3, theta0=110.0, spread=15.0):
    """M chains of four points: bond lengths `bond`, both angles theta0 +-
    spread degrees, uniform dihedrals; each centred on the origin."""
    u1 = _unit(rng.normal(size=(M, 3)))
    pts = [np.zeros((M, 3)), bond * u1]
    prev = u1
    ref = _unit(np.cross(u1, rng.normal(size=(M, 3))))
    for _ in range(2):
        th = np.radians(theta0 + rng.uniform(-spread, spread, size=(M, 1)))
        phi = rng.uniform(0, 2 * math.pi, size=(M, 1))
        b = np.cross(prev, ref)
        # the direction making the angle th with -prev at the last point
        d = -np.cos(th) * prev + np.sin(th) * (np.cos(phi) * ref + np.sin(phi) * b)
        pts.append(pts[-1] + bond * d)
        ref = _unit(np.cross(d, prev))
        prev = d
    P = np.stack(pts, 1)
    return P - P.mean(1, keepdims=True)


def molecules(n=8, seed=8, delta=0.0, dt=0.002):
    rng = _rng(seed)
    nmol = n ** 3
    spacing = 1.6
    L = n * spacing
    cells = max(3, int(math.floor(L / 1.0)))
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    centres = (g + 0.5) * spacing + rng.uniform(-0.15, 0.15, size=(nmol, 3))
    x = ((centres[:, None, :] + _chains(rng, nmol)) % L).reshape(-1, 3)
    N = 4 * nmol
    v = rng.normal(0.0, 0.5, size=(N, 3))
    base = 4 * np.arange(nmol, dtype=np.int32)
    sig, eps = 0.25, 0.5
    A, B = 4 * eps * sig ** 12, 4 * eps * sig ** 6
    th0 = math.radians(110.0)
    return {
        "name": "molecules",
        "dim": [L, L, L], "cells": [cells] * 3, "cutoff": 1.0, "dt": dt,
        "types": [{"mass": 1.0, "radius": 0.1, "dynamics": 0}],
        "x": x, "v": v, "type": np.zeros(N, np.int32),
        "pots": [("lj126", (0.1, 1.0, A, B, 1e-4), 0, 0)],
        "bonds": {"i": np.concatenate([base, base + 1, base + 2]),
                  "j": np.concatenate([base + 1, base + 2, base + 3]),
                  "pot": ("harmonic", (0.0, 0.6, 100.0, 0.3, 1e-4)), "dissociation": 1e30},
        "angles": {"i": np.concatenate([base, base + 1]), "j": np.concatenate([base + 1, base + 2]),
                   "k": np.concatenate([base + 2, base + 3]),
                   "pot": ("harmonic_angle", (th0 - math.radians(60), th0 + math.radians(60), 20.0, th0, 1e-4))},
        # a delta off the multiples of pi puts sqrt(1 - c^2) into the table,
        # whose end slopes the builder only meets at a coarser tolerance
        "dihedrals": {"i": base, "j": base + 1, "k": base + 2, "l": base + 3,
                      "pot": ("harmonic_dihedral", (1.0, 3, delta, 1e-4 if math.fmod(delta, math.pi) == 0 else 1e-2))},
        "exclusions": {"i": np.concatenate([base, base + 1, base + 2, base, base + 1]),
                       "j": np.concatenate([base + 1, base + 2, base + 3, base + 2, base + 3])},
    }


def channel(N=1048576, seed=9):
    rng = _rng(seed)
    density = 15 ** 3 / (15.0 * 12.0 * 10.0)
    s = (N / density / 1800.0) ** (1.0 / 3.0)
    dim = [15.0 * s, 12.0 * s, 10.0 * s]
    cells = [max(3, int(math.floor(d))) for d in dim]
    x = rng.random((N, 3)) * dim
    return {
        "name": "channel",
        "dim": dim, "cells": cells, "cutoff": 1.0, "dt": 0.1,
        "types": [{"mass": 10.0, "radius": 0.3, "dynamics": 0}],
        "x": x, "v": np.zeros((N, 3)), "type": np.zeros(N, np.int32),
        "pots": [("dpd", (1.0, 1.0, 1.5, 1.0, 0), 0, 0)],
        "bonds": None,
        "boundaries": {"kinds": [2, 2, 2, 2, 1, 1]},
        "single_body": {0: (3, np.array([0.01, 0.0, 0.0], np.float32))},
    }


def species_fluid(N=1048576, seed=10):
    w = dpd_fluid(N=N, sigma=3.0, seed=seed)
    rng = _rng(seed + 1)
    w["name"] = "species"
    w["types"] = [dict(w["types"][0]), dict(w["types"][0])]
    w["type"] = (rng.random(N) < 0.5).astype(np.int32)
    dpd = w["pots"][0]
    w["pots"] = [dpd[:2] + (0, 0), dpd[:2] + (0, 1), dpd[:2] + (1, 1)]
    w["species"] = (rng.random((N, 3)) * [2.0, 1.0, 3.0]).astype(np.float32)
    w["fluxes"] = [(0, 0, 0, 0, 0, 1.0, 0.0, 0.0), (0, 1, 1, 0, 0, 1.0, 0.0, 0.0),
                   (0, 0, 1, 0, 0, 0.5, 0.01, 0.0), (1, 0, 1, 1, 1, 0.3, 0.0, 0.5),
                   (2, 1, 0, 2, 2, 0.2, 0.0, 2.0)]
    return w


def cell_sort_replica(N, seed, kT=0.1, dt=0.01, density=0.8):
    rng = _rng(seed)
    L = (N / density) ** (1.0 / 3.0)
    cells = max(3, int(math.floor(L / 2.5)))
    x = rng.random((N, 3)) * L
    t = np.zeros(N, np.int32)
    t[N // 2:] = 1
    rng.shuffle(t)
    return {"dim": [L, L, L], "cells": [cells] * 3, "x": x, "type": t}


def cell_sort_params(kT=0.1, dt=0.01, mass=1.0):
    # Per-component variance of the random force 2 kT m / dt; the magnitude
    # of MxRandomVector is N(0, std) over a uniform direction, so each
    # component has variance std^2 / 3.
    std = math.sqrt(3.0 * 2.0 * kT * mass / dt)
    return {
        "cutoff": 2.5, "dt": dt,
        "types": [{"mass": mass, "radius": 1.0, "dynamics": 1}, {"mass": mass, "radius": 1.0, "dynamics": 1}],
        "pots": [("morse", (3.0, 6.0, 1.0, 1e-4, 2.5, 1e-3), 0, 0),
                 ("morse", (3.0, 6.0, 1.0, 1e-4, 2.5, 1e-3), 1, 1),
                 ("morse", (1.0, 6.0, 1.0, 1e-4, 2.5, 1e-3), 0, 1)],
        "random": {"mean": 0.0, "std": std, "duration": dt},
    }


def cell_sort(N=4000, seed=5, kT=0.1, dt=0.01):
    """Two-type cell sorting: Morse adhesion plus the single-body random force
    on every type, overdamped."""
    w = cell_sort_replica(N, seed, kT=kT, dt=dt)
    w.update(cell_sort_params(kT=kT, dt=dt))
    w["name"] = "cell_sort"
    w["v"] = np.zeros_like(w["x"])
    w["bonds"] = None
    return w


def adjoint_morse(N=65536, L=40.0, seed=6, min_dist=None):
    """Two-type overdamped Morse system for the adjoint: uniform start in a
    periodic box, Morse d = 3 (AA, BB) and 1 (AB), a = 6, r0 = 1, cutoff
    2.5, dt 0.01; theta rows (d, a, r0) for AA, AB, BB."""
    rng = _rng(seed)
    if min_dist:
        x = _min_dist_fill(rng, N, L, min_dist)
    else:
        x = rng.random((N, 3)) * L
    t = np.zeros(N, np.int32)
    t[N // 2:] = 1
    rng.shuffle(t)
    theta = np.array([[3.0, 6.0, 1.0], [1.0, 6.0, 1.0], [3.0, 6.0, 1.0]])
    return {"name": "adjoint_morse", "dim": [L, L, L], "cutoff": 2.5, "dt": 0.01, "x": x, "type": t,
            "mass": (1.0, 1.0), "radius": (1.0, 1.0), "theta": theta, "smin": 1e-4, "smax": 2.5}


def adjoint_mixed(N=4096, L=16.0, seed=8, min_dist=None, bond_r=1.3):
    """The adjoint's widened system: two types with Morse AA and BB (d = 3,
    a = 6, r0 = 1 on [1e-4, 2.5]) and Lennard-Jones AB (A = 0.987,
    B = 1.987: sigma 0.89, epsilon 1, minimum near r = 1, on [0.5, 2.5]),
    plus harmonic bonds (K = 10, r0 = 1 on [0.2, 2.5]) between the pairs
    closer than bond_r at the start; theta rows AA, AB, BB, bond."""
    from scipy.spatial import cKDTree
    w = adjoint_morse(N=N, L=L, seed=seed, min_dist=min_dist)
    tree = cKDTree(np.mod(w["x"], L), boxsize=L)
    pairs = tree.query_pairs(bond_r, output_type="ndarray")
    w["name"] = "adjoint_mixed"
    w["theta"] = np.array([[3.0, 6.0, 1.0], [0.987, 1.987, 0.0], [3.0, 6.0, 1.0], [10.0, 1.0, 0.0]])
    w["kinds"] = {(0, 1): (1, 0.5, 2.5)}
    w["bonds"] = (pairs[:, 0].astype(np.int32), pairs[:, 1].astype(np.int32), 0.2, 2.5)
    return w


def rdf_target(w, nbins=64, rmax=2.5, seed=7):
    """A seeded target g(r): the soft g(r) of an independent uniform
    configuration of the same size and composition, on the host."""
    rng = _rng(seed)
    L = np.asarray(w["dim"], np.float64)
    N = len(w["x"])
    x = rng.random((N, 3)) * L
    t = np.asarray(w["type"])
    nt = int(t.max()) + 1
    np_ = nt * (nt + 1) // 2
    dr = rmax / nbins
    H = np.zeros((np_, nbins))
    cells = np.maximum(3, np.floor(L / rmax)).astype(int)
    c = np.minimum((x / (L / cells)).astype(int), cells - 1)
    lin = (c[:, 0] * cells[1] + c[:, 1]) * cells[2] + c[:, 2]
    order = np.argsort(lin, kind="stable")
    bounds = np.searchsorted(lin[order], np.arange(np.prod(cells) + 1))
    pidx = {}
    k = 0
    for a in range(nt):
        for b in range(a, nt):
            pidx[(a, b)] = pidx[(b, a)] = k
            k += 1
    for cx in range(cells[0]):
        for cy in range(cells[1]):
            for cz in range(cells[2]):
                me = order[bounds[(cx * cells[1] + cy) * cells[2] + cz]:bounds[(cx * cells[1] + cy) * cells[2] + cz + 1]]
                if len(me) == 0:
                    continue
                nb = []
                for ox in (-1, 0, 1):
                    for oy in (-1, 0, 1):
                        for oz in (-1, 0, 1):
                            q = ((cx + ox) % cells[0] * cells[1] + (cy + oy) % cells[1]) * cells[2] + (cz + oz) % cells[2]
                            nb.append(order[bounds[q]:bounds[q + 1]])
                nb = np.unique(np.concatenate(nb))
                d = x[me][:, None, :] - x[nb][None, :, :]
                d -= L * np.round(d / L)
                r = np.sqrt((d * d).sum(-1))
                ii, jj = np.nonzero((r < rmax) & (me[:, None] < nb[None, :]))
                rr = r[ii, jj]
                pp = np.array([pidx[(int(t[me[a]]), int(t[nb[b]]))] for a, b in zip(ii, jj)], dtype=int)
                qv = rr / dr - 0.5
                k0 = np.floor(qv).astype(int)
                for off in (0, 1):
                    kk = k0 + off
                    tt = qv - kk
                    ok = (kk >= 0) & (kk < nbins) & (np.abs(tt) < 1)
                    np.add.at(H, (pp[ok], kk[ok]), (1 - tt[ok] ** 2) ** 2)
    n = np.bincount(t, minlength=nt).astype(np.float64)
    pairs = []
    for a in range(nt):
        for b in range(a, nt):
            pairs.append(n[a] * (n[a] - 1) / 2 if a == b else n[a] * n[b])
    r = (np.arange(nbins) + 0.5) * dr
    ideal = np.asarray(pairs)[:, None] * 4 * np.pi * r[None, :] ** 2 * dr / np.prod(L)
    return H / ideal


def ensemble_sizes(R=256, lo=500, hi=20000, seed=4):
    rng = _rng(seed)
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=R)).astype(np.int64)


def ensemble(R=256, seed=4, kT=0.1, dt=0.01, lo=500, hi=20000):
    """R two-type cell-sorting replicas, sizes log-uniform in [lo, hi]."""
    sizes = ensemble_sizes(R, lo=lo, hi=hi, seed=seed)
    reps = [cell_sort_replica(int(n), seed=1000 + i, kT=kT, dt=dt) for i, n in enumerate(sizes)]
    e = {"name": "ensemble", "replicas": reps}
    e.update(cell_sort_params(kT=kT, dt=dt))
    return e


def build_ensemble(m, e, seed=None):
    """One engine holding every replica of ensemble e."""
    u = m.Universe(cutoff=e["cutoff"], dt=e["dt"], replicas=[(r["dim"], r["cells"]) for r in e["replicas"]])
    for t in e["types"]:
        u.add_type(t["mass"], t["radius"], t["dynamics"])
    pots = []
    for spec in e["pots"]:
        p = make_potential(m, spec)
        u.bind(p, spec[2], spec[3])
        pots.append(p)
    if seed is not None:
        u.set_seed(seed)
    rf = e.get("random")
    if rf:
        for t in range(len(e["types"])):
            u.add_random_force(t, rf["mean"], rf["std"], rf["duration"])
    x = np.concatenate([r["x"] for r in e["replicas"]])
    t = np.concatenate([r["type"] for r in e["replicas"]])
    rep = np.concatenate([np.full(len(r["x"]), i, np.int32) for i, r in enumerate(e["replicas"])])
    u.load(x, None, t, replica=rep)
    u._keep = pots
    return u


def boundary_arrays(b):
    """(kinds[6], velocity[6x3], restore[6], radius[6]) from a workload's
    "boundaries" entry: {"kinds": [...], "velocity": ..., "restore": ...,
    "radius": ...} over the sides left, right, front, back, bottom, top."""
    kinds = np.asarray(b["kinds"], np.int32)
    vel = np.asarray(b.get("velocity", np.zeros((6, 3))), np.float32).reshape(6, 3)
    res = np.asarray(b.get("restore", np.ones(6)), np.float32)
    rad = np.asarray(b.get("radius", np.zeros(6)), np.float32)
    return kinds, vel, res, rad


def make_potential(m, spec):
    """Builds a Potential from a (kind, args) spec with the given module."""
    kind, args = spec[0], spec[1]
    if kind == "lj126":
        return m.Potential.lennard_jones_12_6(*args)
    if kind == "morse":
        d, a, r0, lo, hi, tol = args
        return m.Potential.morse(d=d, a=a, r0=r0, min=lo, max=hi, tol=tol)
    if kind == "harmonic":
        lo, hi, k, r0, tol = args
        return m.Potential.harmonic(k, r0, min=lo, max=hi, tol=tol)
    if kind == "harmonic_angle":
        lo, hi, k, th0, tol = args
        return m.Potential.harmonic_angle(k, th0, min=lo, max=hi, tol=tol)
    if kind == "harmonic_dihedral":
        k, n, delta, tol = args
        return m.Potential.harmonic_dihedral(k, n, delta, tol=tol)
    if kind == "dpd":
        al, g, s, c, sh = args
        return m.Potential.dpd(alpha=al, gamma=g, sigma=s, cutoff=c, shifted=bool(sh))
    raise ValueError(kind)


def build_universe(m, w, seed=None):
    """Creates and loads a Universe for workload dict w."""
    u = m.Universe(dim=w["dim"], cells=w["cells"], cutoff=w["cutoff"], dt=w["dt"])
    for t in w["types"]:
        u.add_type(t["mass"], t["radius"], t["dynamics"])
    pots = []
    for spec in w["pots"]:
        p = make_potential(m, spec)
        u.bind(p, spec[2], spec[3])
        pots.append(p)
    if w.get("boundaries"):
        k, vel, res, rad = boundary_arrays(w["boundaries"])
        u.set_boundaries([int(x) for x in k], vel, res, rad)
    for side, t, spec in w.get("wall_pots") or ():
        p = make_potential(m, spec)
        u.set_boundary_potential(side, t, p)
        pots.append(p)
    if seed is not None:
        u.set_seed(seed)
    if w.get("random"):
        rf = w["random"]
        for t in range(len(w["types"])):
            u.add_random_force(t, rf["mean"], rf["std"], rf["duration"])
    if w.get("dpd_noise"):
        u.set_dpd_noise(1)
    if w.get("integrator"):
        u.set_integrator(w["integrator"])
    # single_body: {type: (kind, params)}, kinds 2 berendsen (tau), 3 constant
    # (F), 4 friction (mean, std, duration, coef)
    for t, (kind, params) in (w.get("single_body") or {}).items():
        if kind == 2:
            u.add_berendsen(t, params[0])
        elif kind == 3:
            u.add_constant_force(t, params)
        elif kind == 4:
            u.add_friction(t, params[3], params[0], params[1], params[2])
    for t, e in (w.get("target_energy") or {}).items():
        u.set_target_energy(t, e)
    u.load(w["x"], w["v"], w["type"], flags=w.get("flags"))
    if w.get("bonds"):
        b = w["bonds"]
        bp = make_potential(m, b["pot"])
        u.set_bonds(b["i"], b["j"], bp, b["dissociation"])
        pots.append(bp)
    for t, mask in (w.get("species_const") or {}).items():
        u.set_species(t, mask)
    for fx in w.get("fluxes") or ():
        u.add_flux(*fx)
    if w.get("species") is not None:
        u.set_species_values(w["species"])
    if w.get("exclusions"):
        u.set_exclusions(w["exclusions"]["i"], w["exclusions"]["j"])
    if w.get("angles"):
        a = w["angles"]
        ap = make_potential(m, a["pot"])
        u.set_angles(a["i"], a["j"], a["k"], ap)
        pots.append(ap)
    if w.get("dihedrals"):
        d = w["dihedrals"]
        dp = make_potential(m, d["pot"])
        u.set_dihedrals(d["i"], d["j"], d["k"], d["l"], dp)
        pots.append(dp)
    u._keep = pots
    return u
