"""Edge cases of the step against the oracle: a lone particle, a single pair
(inside one box and across the periodic seam), a sparse box where nearly
every cell is empty, a clustered fluid whose cells hold hundreds of
particles (bricks overflow LDS and run the global-memory kernel), particles
lying exactly on cell faces or loaded just below one (the first step's
local shuffle moves them on), a particle whose advance ends on a face (x + h
rounds to h, and the next step starts from the neighbouring cell, under
Euler and Runge-Kutta 4), bonded chains across the periodic seam, an
ensemble with one-particle, empty and two-particle replicas, particles
loaded just below a wall or just below zero, frozen-axis flags, and loads
the engine refuses.

Bars as in test_gpu_parity: neighbour counts bit for bit; forces within 1e-5
of the FP32 oracle per particle (floor 0.1 rms |F|; all zero where the
oracle's are); the number density within 1e-3 of the largest (the oracle
keeps the reference's approximate rsqrt); a five-step relay (the GPU's
forces fed to the oracle's advance) bit for bit in positions, velocities
and cells."""
import math

import numpy as np
import pytest

from mechanica_amd import workloads as W

pytestmark = pytest.mark.gpu


def _dpd(x, L, cells, sigma=0.0, seed=1):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = len(x)
    return {
        "name": "dpd_edge",
        "dim": [L, L, L], "cells": [cells] * 3, "cutoff": 1.0, "dt": 0.01,
        "types": [{"mass": 1.0, "radius": 1.0, "dynamics": 0}],
        "x": np.asarray(x, np.float64), "v": rng.normal(0.0, 1.0, size=(n, 3)),
        "type": np.zeros(n, np.int32),
        "pots": [("dpd", (25.0, 4.5, sigma, 1.0, 0), 0, 0)],
        "bonds": None,
    }


def _lj(x, L, cells, seed=1):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = len(x)
    return {
        "name": "lj_edge",
        "dim": [L, L, L], "cells": [cells] * 3, "cutoff": 1.0, "dt": 0.01,
        "types": [{"mass": 39.4, "radius": 0.1, "dynamics": 0}],
        "x": np.asarray(x, np.float64), "v": rng.normal(0.0, math.sqrt(1.0 / 39.4), size=(n, 3)),
        "type": np.zeros(n, np.int32),
        "pots": [("lj126", (0.275, 1.0, W.LJ_A, W.LJ_B, 1e-3), 0, 0)],
        "bonds": None,
    }


def check(m, oracle_mod, w, relay_steps=5):
    print("case:", w["name"], "N=%d" % len(w["x"]), flush=True)
    u = W.build_universe(m, w, seed=5)
    o = oracle_mod.from_workload(w, m, seed=5)
    counts = u.neighbour_counts()
    u.step(1)
    g = u.state()
    o.force()
    r = o.state()
    assert np.array_equal(counts, r["ncount"])
    fr = r["f"].astype(np.float64)
    fg = g["f"].astype(np.float64)
    n = np.linalg.norm(fr, axis=1)
    rms = math.sqrt((n ** 2).mean())
    if rms == 0.0:
        assert not fg.any()
    else:
        rel = (np.linalg.norm(fg - fr, axis=1) / np.maximum(n, 0.1 * rms)).max()
        print("max rel force diff %.3e" % rel)
        assert rel <= 1e-5
    nd_g, nd_r = u.number_density(), o.number_density()
    assert np.abs(nd_g - nd_r).max() <= 1e-3 * np.abs(nd_r).max()
    u2 = W.build_universe(m, w, seed=5)
    o2 = oracle_mod.from_workload(w, m, seed=5)
    for k in range(relay_steps):
        u2.step(1)
        s = u2.state()
        o2.step_with_forces(s["f"])
        t = o2.state()
        assert np.array_equal(s["cell"], t["cell"]), k
        assert np.array_equal(s["x"].view(np.uint32), t["x"].view(np.uint32)), k
        assert np.array_equal(s["v"].view(np.uint32), t["v"].view(np.uint32)), k
    return counts, g


def test_lone_particle(m, oracle_mod):
    counts, g = check(m, oracle_mod, _dpd([[2.5, 1.25, 4.75]], 5.0, 5))
    assert counts.tolist() == [0]
    assert not g["f"].any()


@pytest.mark.parametrize("where", ["inside", "across_seam"])
def test_single_pair(m, oracle_mod, where):
    """One LJ pair 0.6 apart: in neighbouring cells, or across the periodic
    boundary (x = 0.2 and L - 0.4).  Each end computes the pair in the
    reference orientation, so the forces are exact negatives."""
    x = [[1.7, 2.5, 2.5], [2.3, 2.5, 2.5]] if where == "inside" else [[0.2, 2.5, 2.5], [4.6, 2.5, 2.5]]
    counts, g = check(m, oracle_mod, _lj(x, 5.0, 5))
    assert counts.tolist() == [1, 1]
    f = g["f"]
    assert np.array_equal(f[0], -f[1])
    assert np.abs(f[0]).max() > 0


def test_sparse_box(m, oracle_mod):
    """60 particles in 12^3 cells: nearly every cell and most bricks are
    empty, and most particles have no neighbour."""
    rng = np.random.Generator(np.random.PCG64(31))
    check(m, oracle_mod, _dpd(rng.random((60, 3)) * 12.0, 12.0, 12))


def test_clustered_cells(m, oracle_mod):
    """1000 DPD particles in a 2 x 2 x 2-cell block of a 10^3-cell box:
    about 125 particles per occupied cell and several hundred neighbours
    per particle, so the bricks around the block overflow LDS (the
    global-memory kernel computes them) while the rest of the box is
    empty."""
    rng = np.random.Generator(np.random.PCG64(32))
    x = 4.0 + 2.0 * rng.random((1000, 3))
    counts, _ = check(m, oracle_mod, _dpd(x, 10.0, 10))
    assert counts.max() > 300


def test_particles_on_cell_faces(m, oracle_mod):
    """LJ on a grid of spacing 0.5 in unit cells: every particle lies on a
    cell face, edge or corner (cell-local coordinate 0), with half of them
    jittered off the grid by up to 0.03, and four 1e-9 below the box's upper
    faces, at least 0.3 from every other particle (no pair comes closer than
    the table's 0.275)."""
    rng = np.random.Generator(np.random.PCG64(33))
    idx = np.arange(12) * 0.5
    x = np.stack(np.meshgrid(idx, idx, idx, indexing="ij"), -1).reshape(-1, 3)
    move = rng.random(len(x)) < 0.5
    x[move] += rng.uniform(-0.03, 0.03, size=(move.sum(), 3))
    x %= 6.0
    x[:4] = [[6.0 - 1e-9, 0.25, 0.25], [0.75, 6.0 - 1e-9, 0.75], [1.75, 1.75, 6.0 - 1e-9],
             [6.0 - 1e-9, 6.0 - 1e-9, 6.0 - 1e-9]]
    check(m, oracle_mod, _lj(x, 6.0, 6))


def test_advance_lands_on_face(m, oracle_mod):
    """A particle 1e-10 above a cell's lower face moving down by 2e-10 a
    step: the advance puts it at -1e-10, moves it to the lower cell at
    -1e-10 + h, which rounds to h, and the reference's next step begins by
    moving it back (space_shuffle_local) before it advances again.  A second
    particle does the same across the periodic seam, a third on two axes."""
    x = [[2.0 + 1e-10, 2.5, 2.5], [1e-10, 0.5, 4.5], [4.5, 3.0 + 1e-10, 1.0 + 1e-10]]
    w = _dpd(x, 5.0, 5)
    w["v"] = np.array([[-2e-8, 0.0, 0.0], [-2e-8, 0.0, 0.0], [0.0, -2e-8, -2e-8]])
    counts, g = check(m, oracle_mod, w)
    assert counts.tolist() == [0, 0, 0]


def test_advance_lands_on_face_rk4(m, oracle_mod):
    """The same particles under Runge-Kutta 4 (each stage's move, and the
    local shuffle that begins each force stage, carry p0 along), and two
    loaded 1e-9 below zero, whose load shuffle leaves them on the face of
    the last cell.  With no pair forces the stages are exact, so the states
    match the oracle's RK4 bit for bit, step after step."""
    x = [[2.0 + 1e-10, 2.5, 2.5], [1e-10, 0.5, 4.5], [4.5, 3.0 + 1e-10, 1.0 + 1e-10], [-1e-9, 2.5, 2.5],
         [3.5, -1e-9, 0.5]]
    w = _dpd(x, 5.0, 5)
    w["v"] = np.array([[-2e-8, 0.0, 0.0], [-2e-8, 0.0, 0.0], [0.0, -2e-8, -2e-8], [1e-8, 0.0, 0.0],
                       [0.0, -1e-8, 0.0]])
    w["integrator"] = 1
    print("case: rk4 face landing", flush=True)
    u = W.build_universe(m, w, seed=5)
    o = oracle_mod.from_workload(w, m, seed=5)
    for k in range(4):
        u.step(1)
        o.step(1)
        s, t = u.state(), o.state()
        assert np.array_equal(s["cell"], t["cell"]), k
        assert np.array_equal(s["x"].view(np.uint32), t["x"].view(np.uint32)), k
        assert np.array_equal(s["v"].view(np.uint32), t["v"].view(np.uint32)), k


def test_bonded_terms_across_the_seam(m, oracle_mod):
    """Two four-bead chains (the molecules' bonds, angles, dihedral,
    exclusions and LJ) that cross the periodic boundary: one across the x
    face, one through the corner where all three axes wrap, so bonds, angles
    and dihedrals take the cell-offset folding across the seam on every
    axis."""
    th0 = math.radians(110.0)
    sig, eps = 0.25, 0.5
    A, B = 4 * eps * sig ** 12, 4 * eps * sig ** 6
    x = np.array([[5.70, 2.50, 2.50], [5.95, 2.60, 2.50], [0.20, 2.50, 2.60], [0.45, 2.60, 2.75],
                  [5.80, 5.85, 5.90], [0.05, 5.95, 0.10], [0.15, 0.20, 0.05], [0.40, 0.30, 0.25]])
    rng = np.random.Generator(np.random.PCG64(34))
    base = np.array([0, 4], np.int32)
    w = {
        "name": "seam_chains", "dim": [6.0, 6.0, 6.0], "cells": [6, 6, 6], "cutoff": 1.0, "dt": 0.002,
        "types": [{"mass": 1.0, "radius": 0.1, "dynamics": 0}],
        "x": x, "v": rng.normal(0.0, 0.5, size=(8, 3)), "type": np.zeros(8, np.int32),
        "pots": [("lj126", (0.1, 1.0, A, B, 1e-4), 0, 0)],
        "bonds": {"i": np.concatenate([base, base + 1, base + 2]), "j": np.concatenate([base + 1, base + 2, base + 3]),
                  "pot": ("harmonic", (0.0, 0.6, 100.0, 0.3, 1e-4)), "dissociation": 1e30},
        "angles": {"i": np.concatenate([base, base + 1]), "j": np.concatenate([base + 1, base + 2]),
                   "k": np.concatenate([base + 2, base + 3]),
                   "pot": ("harmonic_angle", (th0 - math.radians(60), th0 + math.radians(60), 20.0, th0, 1e-4))},
        "dihedrals": {"i": base, "j": base + 1, "k": base + 2, "l": base + 3,
                      "pot": ("harmonic_dihedral", (1.0, 3, 0.0, 1e-4))},
        "exclusions": {"i": np.concatenate([base, base + 1, base + 2, base, base + 1]),
                       "j": np.concatenate([base + 1, base + 2, base + 3, base + 2, base + 3])},
    }
    check(m, oracle_mod, w)


def test_ensemble_tiny_and_empty_replicas(m, oracle_mod):
    """One engine with four cell-sorting replicas (Morse, overdamped, boxes
    of 3^3 cells of width 2.5): one particle, none, a pair 1.2 apart, and
    300 particles.  Each non-empty replica against its own oracle: neighbour
    counts, forces, then a five-step relay bit for bit; the empty replica's
    observables stay zero."""
    params = W.cell_sort_params()
    params.pop("random")
    rng = np.random.Generator(np.random.PCG64(35))
    pts = [np.array([[3.0, 4.0, 5.0]]), np.zeros((0, 3)), np.array([[1.0, 3.7, 3.7], [2.2, 3.7, 3.7]]),
           rng.random((300, 3)) * 7.5]
    reps = [{"dim": [7.5] * 3, "cells": [3] * 3, "x": x, "type": (np.arange(len(x)) % 2).astype(np.int32)}
            for x in pts]
    e = {"name": "ensemble_edges", "replicas": reps}
    e.update(params)
    print("case: ensemble with replicas of", [len(x) for x in pts], "particles", flush=True)

    def replica(r):
        w = dict(reps[r])
        for k in ("cutoff", "dt", "types", "pots"):
            w[k] = e[k]
        w["name"] = "cell_sort"
        w["v"] = np.zeros_like(reps[r]["x"])
        w["bonds"] = None
        return w

    bounds = np.cumsum([0] + [len(x) for x in pts])
    cbase = np.cumsum([0] + [27] * len(pts))
    u = W.build_ensemble(m, e, seed=12)
    counts = u.neighbour_counts()
    u.step(1)
    g = u.state()
    for r in range(len(pts)):
        if len(pts[r]) == 0:
            ob = u.observables(r)
            assert ob["epot_nonbond"] == 0.0 and ob["ekin"] == 0.0
            continue
        sl = slice(bounds[r], bounds[r + 1])
        o = oracle_mod.from_workload(replica(r), m)
        o.force()
        t = o.state()
        assert np.array_equal(counts[sl], t["ncount"]), r
        fr = t["f"].astype(np.float64)
        n = np.linalg.norm(fr, axis=1)
        rms = math.sqrt((n ** 2).mean())
        fg = g["f"][sl].astype(np.float64)
        if rms == 0.0:
            assert not fg.any(), r
        else:
            assert (np.linalg.norm(fg - fr, axis=1) / np.maximum(n, 0.1 * rms)).max() <= 1e-5, r
    u2 = W.build_ensemble(m, e, seed=12)
    oracles = {r: oracle_mod.from_workload(replica(r), m) for r in range(len(pts)) if len(pts[r])}
    for k in range(5):
        u2.step(1)
        s = u2.state()
        for r, o in oracles.items():
            sl = slice(bounds[r], bounds[r + 1])
            o.step_with_forces(s["f"][sl])
            t = o.state()
            assert np.array_equal(s["cell"][sl] - cbase[r], t["cell"]), (k, r)
            assert np.array_equal(s["x"][sl].view(np.uint32), t["x"].view(np.uint32)), (k, r)


def test_wall_face_load(m, oracle_mod):
    """DPD between free-slip walls in z (periodic in x and y) with particles
    loaded 1e-9 below the top wall: their local z rounds to the cell width,
    and the reference's local shuffle moves them on modulo the grid, as on a
    periodic axis, into the bottom cell at z = 0 (space.cpp:333-393 wraps
    every axis)."""
    rng = np.random.Generator(np.random.PCG64(36))
    x = rng.random((400, 3)) * 8.0
    x[:3, 2] = 8.0 - 1e-9
    w = _dpd(x, 8.0, 8)
    w["boundaries"] = {"kinds": [2, 2, 2, 2, 4, 4]}
    check(m, oracle_mod, w)


def test_load_just_below_zero(m, oracle_mod):
    """space_addpart truncates (x - origin) / h toward zero, so a coordinate
    in (-h, 0) is accepted into cell 0 with a negative local coordinate, and
    the local shuffle then moves it to the last cell (x + h)."""
    rng = np.random.Generator(np.random.PCG64(38))
    x = rng.random((200, 3)) * 5.0
    x[:3] = [[1.0, -0.1, 1.0], [-0.5, 2.0, 3.0], [-1e-9, -1e-9, 4.5]]
    counts, _ = check(m, oracle_mod, _dpd(x, 5.0, 5))


def test_frozen_flags(m, oracle_mod):
    """MxParticleFlags on a DPD fluid: PARTICLE_FROZEN_X, FROZEN_Y | FROZEN_Z
    and FROZEN (1 << 3 ... 1 << 5, MxParticle.h:58-60) hold those
    coordinates, PARTICLE_BOUND (1 << 2) holds nothing; against the oracle,
    and the frozen coordinates unchanged over five steps."""
    rng = np.random.Generator(np.random.PCG64(39))
    x = rng.random((600, 3)) * 6.0
    w = _dpd(x, 6.0, 6)
    fl = np.zeros(600, np.uint16)
    fl[0:40] = m.FROZEN_X
    fl[40:80] = m.FROZEN_Y | m.FROZEN_Z
    fl[80:120] = m.FROZEN_X | m.FROZEN_Y | m.FROZEN_Z
    fl[120:160] = 1 << 2
    w["flags"] = fl
    u = W.build_universe(m, w, seed=5)
    x0, _ = u.unload()
    check(m, oracle_mod, w)
    u.step(5)
    x1, _ = u.unload()
    for lo, axes in ((0, [0]), (40, [1, 2]), (80, [0, 1, 2])):
        assert np.array_equal(x1[lo:lo + 40, axes], x0[lo:lo + 40, axes]), lo
        moving = [k for k in range(3) if k not in axes]
        if moving:
            assert (x1[lo:lo + 40, moving] != x0[lo:lo + 40, moving]).all(), lo
    assert (x1[120:160] != x0[120:160]).all()


def test_no_slip_kind(m):
    """BOUNDARY_NO_SLIP sides (1 << 4) behave as velocity walls at rest, as
    the reference's configuration makes them: five steps bit for bit."""
    rng = np.random.Generator(np.random.PCG64(40))
    x = rng.random((500, 3)) * 8.0
    states = []
    for kinds in ([2, 2, 2, 2, m.BOUNDARY_NO_SLIP, m.BOUNDARY_NO_SLIP], [2, 2, 2, 2, 1, 1]):
        w = _dpd(x, 8.0, 8)
        w["boundaries"] = {"kinds": kinds}
        u = W.build_universe(m, w, seed=5)
        u.step(5)
        states.append(u.state())
    for k in ("x", "v", "f", "cell"):
        assert np.array_equal(states[0][k].view(np.uint32), states[1][k].view(np.uint32)), k


def test_load_errors(m):
    """Loads the engine refuses, leaving it unloaded: a particle outside the
    box, beyond it or a cell width or more below it (MXG_ERR_SPACE), an id
    given twice and a type out of range
    (MXG_ERR_RANGE), and 65,536 particles in one cell (MXG_ERR_MAXPARTS,
    mdcore's per-cell limit)."""
    def universe():
        u = m.Universe(dim=[5.0, 5.0, 5.0], cells=[5, 5, 5], cutoff=1.0, dt=0.01)
        u.add_type(1.0, 1.0, 0)
        return u

    cases = [
        (-3, dict(x=[[5.5, 1.0, 1.0]])),
        (-3, dict(x=[[1.0, -1.5, 1.0]])),
        (-6, dict(x=[[1.0, 1.0, 1.0], [2.0, 2.0, 2.0]], pid=[3, 3])),
        (-6, dict(x=[[1.0, 1.0, 1.0]], types=[1])),
        (-24, dict(x=np.random.default_rng(37).random((65536, 3)) * 0.9)),
    ]
    for code, kw in cases:
        u = universe()
        x = np.asarray(kw.pop("x"), np.float64)
        with pytest.raises(m.EngineError) as e:
            u.load(x, None, kw.pop("types", np.zeros(len(x), np.int32)), **kw)
        assert e.value.code == code, (code, e.value)
        with pytest.raises(m.EngineError):
            u.step(1)


def test_empty_load_rejected(m):
    u = m.Universe(dim=[5.0, 5.0, 5.0], cells=[5, 5, 5], cutoff=1.0, dt=0.01)
    u.add_type(1.0, 1.0, 0)
    with pytest.raises(m.EngineError) as e:
        u.load(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.int32))
    assert e.value.code == -1  # MXG_ERR_NULL
    with pytest.raises(m.EngineError):
        u.step(1)
