"""Boundary conditions on the GPU (MxBoundaryConditions, boundary_eval.hpp,
space.cpp:700-888) against the oracle: non-periodic axes, free-slip and
velocity walls in the advance, table and DPD wall potentials in the force
phase.

Bars as the other parity tests: forces within 1e-5 relative to the FP32
oracle (floor 0.1 rms), neighbour counts exact, and the advance relay (the
oracle stepping with the GPU's forces, reflections included) bit for bit.
"""
import numpy as np
import pytest

from mechanica_amd import workloads as W

pytestmark = pytest.mark.gpu

FREESLIP, VELOCITY, PERIODIC, POTENTIAL = 4, 1, 2, 8
LJ = ("lj126", (0.1, 1.0, 4 * 0.5 * 0.25 ** 12, 4 * 0.5 * 0.25 ** 6, 1e-4))
DPD = ("dpd", (25.0, 4.5, 0.0, 1.0, 0))
# A soft wall, K (r - 0.6)^2 within 1 of the wall: boundary_potential_eval_ex
# moves any particle with r^2 < a (not a^2) to the distance a, and this
# table's a is 0.
HARMONIC_WALL = ("harmonic", (0.0, 1.0, 10.0, 0.6, 1e-4))


def rel_err(fg, fr):
    n = np.linalg.norm(fr, axis=1)
    return (np.linalg.norm(fg - fr, axis=1) / np.maximum(n, 0.1 * np.sqrt((n ** 2).mean()))).max()


def fluid(kind, kinds, N=6000, L=12.0, seed=21, wall_pots=(), velocity=None):
    rng = np.random.default_rng(seed)
    if kind == "dpd":
        pot, speed, dt = DPD, 1.0, 0.01
    else:
        pot, speed, dt = LJ, 0.5, 0.002
    w = {"name": "walls_" + kind, "dim": [L] * 3, "cells": [int(L)] * 3, "cutoff": 1.0, "dt": dt,
         "types": [{"mass": 1.0, "radius": 0.1, "dynamics": 0}],
         "x": (W._min_dist_fill(rng, N, L, 0.3) if kind == "lj" else rng.random((N, 3)) * L),
         "v": rng.normal(0, speed, (N, 3)), "type": np.zeros(N, np.int32),
         "pots": [pot + (0, 0)], "bonds": None, "boundaries": {"kinds": kinds},
         "wall_pots": [(s, 0, p) for s, p in wall_pots]}
    if velocity is not None:
        w["boundaries"]["velocity"] = velocity
    return w


CASES = {
    "dpd_zwalls": lambda: fluid("dpd", [PERIODIC] * 4 + [FREESLIP | POTENTIAL] * 2,
                                wall_pots=[(4, DPD), (5, DPD)]),
    "lj_box": lambda: fluid("lj", [FREESLIP] * 6, N=1500, L=8.0),
    "lj_box_walls": lambda: fluid("lj", [FREESLIP | POTENTIAL] * 6, N=1500, L=8.0,
                                  wall_pots=[(s, HARMONIC_WALL) for s in range(6)]),
    "dpd_couette": lambda: fluid("dpd", [PERIODIC] * 4 + [VELOCITY] * 2,
                                 velocity=np.array([[0, 0, 0]] * 5 + [[1.0, 0, 0]], np.float64)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_boundaries_match_oracle(m, oracle_mod, case):
    w = CASES[case]()
    print("case: boundaries", case, "N=%d" % len(w["x"]), flush=True)
    u = W.build_universe(m, w, seed=5)
    o = oracle_mod.from_workload(w, m, seed=5)
    counts = u.neighbour_counts()
    u.step(1)
    o.force()
    assert np.array_equal(counts, o.state()["ncount"])
    rel = rel_err(u.state()["f"].astype(np.float64), o.state()["f"].astype(np.float64))
    print("rel(gpu, oracle) = %.3e" % rel)
    assert rel <= 1e-5
    og, orr = u.observables(), o.observables()
    assert abs(og["epot_nonbond"] - orr["epot_nonbond"]) <= 1e-5 * max(1.0, abs(orr["epot_nonbond"]))
    u2 = W.build_universe(m, w, seed=5)
    o2 = oracle_mod.from_workload(w, m, seed=5)
    for k in range(10):
        u2.step(1)
        g = u2.state()
        o2.step_with_forces(g["f"])
        r = o2.state()
        assert np.array_equal(g["cell"], r["cell"]), k
        assert np.array_equal(g["x"].view(np.uint32), r["x"].view(np.uint32)), k
        assert np.array_equal(g["v"].view(np.uint32), r["v"].view(np.uint32)), k


def test_global_fallback_with_walls(m, oracle_mod):
    """pmax = 64 sends nearly every brick to the global-memory kernel."""
    w = CASES["lj_box_walls"]()
    print("case: boundaries global fallback", flush=True)
    u = W.build_universe(m, w, seed=5)
    u.set_brick(2, 2, 2, 0, 64)
    o = oracle_mod.from_workload(w, m, seed=5)
    u.step(1)
    o.force()
    assert rel_err(u.state()["f"].astype(np.float64), o.state()["f"].astype(np.float64)) <= 1e-5


def test_freeslip_box_confines(m):
    w = CASES["lj_box"]()
    print("case: free-slip confinement", flush=True)
    u = W.build_universe(m, w, seed=5)
    u.step(500)
    x, _ = u.unload()
    assert x.min() >= 0.0 and x.max() < 8.0


def test_couette_profile(m):
    """A moving top wall (no-slip bottom at rest) drags the fluid: the mean
    x velocity near the top exceeds that near the bottom."""
    w = CASES["dpd_couette"]()
    print("case: couette", flush=True)
    u = W.build_universe(m, w, seed=5)
    u.step(400)
    x, v = u.unload()
    top = v[x[:, 2] > 10.0, 0].mean()
    bottom = v[x[:, 2] < 2.0, 0].mean()
    print("top %.3f bottom %.3f" % (top, bottom))
    assert top > bottom + 0.1
