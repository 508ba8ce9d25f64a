"""Pins the oracle's FP64 virial W = sum over pairs and bonds of
(x_i - x_j) (x) f_i.  The reference's virial_pair never sets its force
(metrics.cpp:183-246), so the definition is pinned by identities rather than
by reference output:

  * in a box with walls on every side (no pair crosses a periodic face)
    W equals sum_i x_i (x) F_i over absolute positions and total forces;
  * for one pair it is r (x) f(r) with f the pair force along r;
  * in a periodic box it does not change when every particle is translated
    (wrapped) by the same vector.
"""
import numpy as np

from mechanica_amd import workloads as W

LJ = ("lj126", (0.1, 1.0, 4 * 0.5 * 0.25 ** 12, 4 * 0.5 * 0.25 ** 6, 1e-4), 0, 0)


def system(x, kinds=None, L=8.0, bonds=None, pots=(LJ,)):
    x = np.asarray(x, np.float64)
    N = len(x)
    w = {"name": "virial", "dim": [L] * 3, "cells": [int(L)] * 3, "cutoff": 1.0, "dt": 0.002,
         "types": [{"mass": 1.0, "radius": 0.1, "dynamics": 0}],
         "x": x, "v": np.zeros((N, 3)), "type": np.zeros(N, np.int32), "pots": list(pots), "bonds": bonds}
    if kinds is not None:
        w["boundaries"] = {"kinds": kinds}
    return w


def test_walled_box_equals_position_moment(m, oracle_mod):
    rng = np.random.default_rng(7)
    N = 800
    x = W._min_dist_fill(rng, N, 8.0, 0.3)
    # bonds to the nearest particle (bond_eval folds cell offsets to one
    # cell, so bonded particles sit in neighbouring cells)
    i = np.arange(200, dtype=np.int32)
    d = np.linalg.norm(x[:200, None, :] - x[None, :, :], axis=2)
    d[i, i] = np.inf
    j = d.argmin(1).astype(np.int32)
    bonds = {"i": i, "j": j, "pot": ("harmonic", (0.0, 2.0, 5.0, 0.5, 1e-4)), "dissociation": 1e30}
    o = oracle_mod.from_workload(system(x, kinds=[4] * 6, bonds=bonds), m)
    F, _ = o.forces64()
    Wv = o.virial64()
    Xg = o.global_positions()
    ref = Xg.T @ F
    assert np.abs(Wv).max() > 1.0
    assert np.allclose(Wv, ref, rtol=0, atol=1e-9 * np.abs(ref).max() * N)


def test_single_pair(m, oracle_mod):
    o = oracle_mod.from_workload(system([[3.5, 3.6, 3.4], np.array([3.5, 3.6, 3.4]) - [0.21, -0.12, 0.05]]), m)
    Xg = o.global_positions()  # the positions as stored (float, cell-relative)
    d = Xg[0] - Xg[1]
    r = np.linalg.norm(d)
    Wv = o.virial64()
    p = W.make_potential(m, LJ[:2])
    F, _ = o.forces64()
    # f_i = F[0]; W = d (x) f_i with f_i parallel to d
    assert np.allclose(np.cross(F[0], d), 0, atol=1e-9 * np.abs(F[0]).max())
    assert np.allclose(Wv, np.outer(d, F[0]), rtol=1e-9, atol=1e-12)
    _, _, f = p.eval_ex(r * r, 0.1, 0.1)
    assert abs(np.trace(Wv) + f * r * r) <= 1e-4 * abs(f * r * r)


def test_periodic_translation_invariance(m, oracle_mod):
    rng = np.random.default_rng(8)
    x = W._min_dist_fill(rng, 600, 8.0, 0.3)
    W0 = oracle_mod.from_workload(system(x), m).virial64()
    W1 = oracle_mod.from_workload(system((x + [2.3, 5.1, 7.7]) % 8.0), m).virial64()
    assert np.allclose(W0, W1, rtol=0, atol=1e-5 * np.abs(W0).max())


def test_dpd_sigma_term_before_first_force(m, oracle_mod):
    """DPD with the reference's deterministic sigma term (sigma * w / sqrt(dt),
    dpd_eval.hpp:24): in a walled box the FP64 virial taken before any force
    phase equals sum_i x_i (x) f_i of the FP32 force phase that follows
    (itself pinned pair by pair in test_oracle.py), and is unchanged by it."""
    rng = np.random.default_rng(9)
    N = 900
    x = W._min_dist_fill(rng, N, 8.0, 0.3)
    w = system(x, kinds=[4] * 6, pots=(("dpd", (25.0, 4.5, 3.0, 1.0, 0), 0, 0),))
    w["v"] = rng.normal(size=(N, 3))
    o = oracle_mod.from_workload(w, m)
    W0 = o.virial64()
    o.force()
    W1 = o.virial64()
    f = o.state()["f"].astype(np.float64)
    ref = o.global_positions().T @ f
    assert np.abs(W0).max() > 1.0
    assert np.array_equal(W0, W1)
    assert np.allclose(W0, ref, rtol=0, atol=1e-5 * np.abs(ref).max())
