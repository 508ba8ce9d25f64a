"""Angles, dihedrals and exclusions on the GPU (angle_eval, angle.cpp:92-354;
dihedral_eval, dihedral.cpp:90-330; exclusion_eval, exclusion.cpp:83-260)
against the oracle, on four-bead molecules with bonds and an LJ fluid
between them.

Bars as the other parity tests: forces within 1e-5 relative to the FP32
oracle (floor 0.1 rms), the three term energies within 1e-6 relative, and
the advance relay (the oracle stepping with the GPU's forces) bit for bit in
positions, velocities and cells.  The random plane of a right angle is
keyed by (angle, time) on both sides, so it is compared like any other
force.
"""
import math

import numpy as np
import pytest

from mechanica_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel_err(fg, fr):
    n = np.linalg.norm(fr, axis=1)
    return (np.linalg.norm(fg - fr, axis=1) / np.maximum(n, 0.1 * np.sqrt((n ** 2).mean()))).max()


@pytest.mark.parametrize("delta", [0.0, math.pi / 4])
def test_terms_match_oracle(m, oracle_mod, delta):
    w = W.molecules(n=11, seed=11, delta=delta)
    print("case: molecules delta %.3f N=%d" % (delta, len(w["x"])), flush=True)
    u = W.build_universe(m, w, seed=5)
    o = oracle_mod.from_workload(w, m, seed=5)
    u.step(1)
    o.force()
    fg = u.state()["f"].astype(np.float64)
    fr = o.state()["f"].astype(np.float64)
    rel = rel_err(fg, fr)
    print("rel(gpu, oracle) = %.3e" % rel)
    assert rel <= 1e-5
    tg, tr = u.term_energies(), o.term_energies()
    for k in tr:
        assert abs(tg[k] - tr[k]) <= 1e-6 * max(1.0, abs(tr[k])), (k, tg[k], tr[k])
    og, orr = u.observables(), o.observables()
    assert abs(og["epot_bond"] - orr["epot_bond"]) <= 1e-6 * max(1.0, abs(orr["epot_bond"]))
    # relay: five steps of the oracle's advance with the GPU's forces
    u2 = W.build_universe(m, w, seed=5)
    o2 = oracle_mod.from_workload(w, m, seed=5)
    for k in range(5):
        u2.step(1)
        g = u2.state()
        o2.step_with_forces(g["f"])
        r = o2.state()
        assert np.array_equal(g["cell"], r["cell"]), k
        assert np.array_equal(g["x"].view(np.uint32), r["x"].view(np.uint32)), k
        assert np.array_equal(g["v"].view(np.uint32), r["v"].view(np.uint32)), k


def test_terms_trajectory(m, oracle_mod):
    """Twenty steps side by side: positions within 1e-4 of the oracle's."""
    w = W.molecules(n=9, seed=12)
    print("case: molecules trajectory", flush=True)
    u = W.build_universe(m, w, seed=5)
    o = oracle_mod.from_workload(w, m, seed=5)
    u.step(20)
    assert o.step(20) == 0
    L = np.asarray(w["dim"])
    d = u.unload()[0] - o.global_positions()
    d -= L * np.round(d / L)
    print("max |dx| %.3e" % np.abs(d).max())
    assert np.abs(d).max() <= 1e-4


def test_right_angles_and_exclusions_only(m, oracle_mod):
    """Angles at exactly 90 degrees (random plane) and exclusions without
    bonds: the term lists without a bond list."""
    n = 64
    rng = np.random.default_rng(13)
    g = np.stack(np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij"), -1).reshape(-1, 3)
    base = g * 2.0 + 0.5
    x = np.concatenate([base + [0.3, 0.0, 0.0], base, base + [0.0, 0.3, 0.0]])
    N = len(x)
    lj = ("lj126", (0.1, 1.0, 4 * 0.5 * 0.25 ** 12, 4 * 0.5 * 0.25 ** 6, 1e-4), 0, 0)
    idx = np.arange(n, dtype=np.int32)
    w = {"name": "right_angles", "dim": [8.0] * 3, "cells": [8] * 3, "cutoff": 1.0, "dt": 0.002,
         "types": [{"mass": 1.0, "radius": 0.1, "dynamics": 0}],
         "x": x, "v": rng.normal(0, 0.1, (N, 3)), "type": np.zeros(N, np.int32), "pots": [lj], "bonds": None,
         "angles": {"i": idx, "j": idx + n, "k": idx + 2 * n,
                    "pot": ("harmonic_angle", (math.radians(50), math.radians(170), 20.0, math.radians(110), 1e-4))},
         "exclusions": {"i": idx, "j": idx + n}}
    print("case: right angles", flush=True)
    u = W.build_universe(m, w, seed=5)
    o = oracle_mod.from_workload(w, m, seed=5)
    u.step(1)
    o.force()
    fg = u.state()["f"].astype(np.float64)
    fr = o.state()["f"].astype(np.float64)
    assert rel_err(fg, fr) <= 1e-5


def test_terms_rk4(m, oracle_mod):
    w = W.molecules(n=8, seed=14)
    w["integrator"] = 1
    print("case: molecules rk4", flush=True)
    u = W.build_universe(m, w, seed=5)
    o = oracle_mod.from_workload(w, m, seed=5)
    u.step(3)
    assert o.step(3) == 0
    L = np.asarray(w["dim"])
    d = u.unload()[0] - o.global_positions()
    d -= L * np.round(d / L)
    assert np.abs(d).max() <= 1e-4


def test_terms_graph_replay_bitwise(m):
    w = W.molecules(n=8, seed=15)
    print("case: molecules graphs", flush=True)
    out = []
    for graphs in (False, True):
        u = W.build_universe(m, w, seed=9)
        u.set_graphs(graphs)
        u.step(3)
        u.state()
        u.step(4)
        s = u.state()
        out.append((s["x"].copy(), s["v"].copy(), s["cell"].copy(), u.term_energies()["epot_angle"]))
    for a, b in zip(out[0][:3], out[1][:3]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert out[0][3] == out[1][3]


def test_terms_replace_and_clear(m, oracle_mod):
    """set_* replaces the previous set; an empty set removes the terms."""
    w = W.molecules(n=7, seed=16)
    u = W.build_universe(m, w, seed=5)
    u.set_angles([], [], [], [])
    u.set_dihedrals([], [], [], [], [])
    u.set_exclusions([], [])
    w2 = dict(w)
    w2["angles"] = w2["dihedrals"] = w2["exclusions"] = None
    o = oracle_mod.from_workload(w2, m, seed=5)
    u.step(1)
    o.force()
    assert rel_err(u.state()["f"].astype(np.float64), o.state()["f"].astype(np.float64)) <= 1e-5
    assert u.term_energies() == {"epot_angle": 0.0, "epot_dihedral": 0.0, "epot_exclusion": 0.0}
