"""Pins the oracle's species transport (flux_eval_ex, flux_eval.hpp:17-81;
MxFluxes_Integrate, Flux.cpp:261-284) with closed forms on resting
particles (no potentials, zero velocity, so nothing moves):

  * Fick between two particles: q = k (s_a - s_b) (1 - r/cutoff)^2 moves
    from a to b, and one Euler step adds dt q;
  * secrete acts only while s_a exceeds the target, uptake only while
    s_b is below it, each scaled by the source's value for uptake;
  * decay takes decay/2 s from each side per pair;
  * Fick without decay conserves the total and flattens the profile;
  * a constant species never changes;
  * a particle that changed cells keeps its accumulated flux for the next
    step (the reference integrates a cell after its own advance, when the
    mover already sits in another cell's incoming list).
"""
import numpy as np

FICK, SECRETE, UPTAKE = 0, 1, 2


def resting(x, types, ntypes=2, L=6.0, dt=0.01, v=None):
    x = np.asarray(x, np.float64)
    N = len(x)
    return {"name": "species", "dim": [L] * 3, "cells": [int(L)] * 3, "cutoff": 1.0, "dt": dt,
            "types": [{"mass": 1.0, "radius": 0.1, "dynamics": 0} for _ in range(ntypes)],
            "x": x, "v": np.zeros((N, 3)) if v is None else np.asarray(v, np.float64),
            "type": np.asarray(types, np.int32), "pots": [], "bonds": None}


def two(kind, sa, sb, k=0.7, decay=0.0, target=0.0):
    w = resting([[2.3, 2.4, 2.5], [2.7, 2.6, 2.45]], [0, 1])
    w["fluxes"] = [(kind, 0, 1, 0, 0, k, decay, target)]
    w["species"] = np.array([[sa, 0, 0, 0], [sb, 0, 0, 0]], np.float32)
    return w


def closed_two(kind, sa, sb, k=0.7, decay=0.0, target=0.0, dt=0.01):
    d = np.array([0.4, 0.2, -0.05])
    r = np.linalg.norm(d)
    term = (1 - r / 1.0) ** 2
    if kind == FICK:
        rate = k * (sa - sb)
    elif kind == SECRETE:
        rate = max(k * (sa - target), 0.0)
    else:
        rate = max(k * (target - sb) * sa, 0.0)
    q = term * rate
    qa = -q - decay / 2 * sa
    qb = q - decay / 2 * sb
    return sa + dt * qa, sb + dt * qb


def run(m, oracle_mod, w, steps=1):
    o = oracle_mod.from_workload(w, m)
    assert o.step(steps) == 0
    return o.species()


def test_fick_two_particles(m, oracle_mod):
    v, q = run(m, oracle_mod, two(FICK, 2.0, 0.5))
    a, b = closed_two(FICK, 2.0, 0.5)
    assert np.allclose(v[:, 0], [a, b], rtol=2e-6)
    assert np.all(q == 0)


def test_secrete_threshold(m, oracle_mod):
    v, _ = run(m, oracle_mod, two(SECRETE, 2.0, 0.5, target=1.0))
    assert np.allclose(v[:, 0], closed_two(SECRETE, 2.0, 0.5, target=1.0), rtol=2e-6)
    v, _ = run(m, oracle_mod, two(SECRETE, 0.8, 0.5, target=1.0))
    assert np.array_equal(v[:, 0], np.float32([0.8, 0.5]))


def test_uptake(m, oracle_mod):
    v, _ = run(m, oracle_mod, two(UPTAKE, 1.5, 0.2, target=1.0))
    assert np.allclose(v[:, 0], closed_two(UPTAKE, 1.5, 0.2, target=1.0), rtol=2e-6)
    v, _ = run(m, oracle_mod, two(UPTAKE, 1.5, 1.2, target=1.0))
    assert np.array_equal(v[:, 0], np.float32([1.5, 1.2]))


def test_decay(m, oracle_mod):
    v, _ = run(m, oracle_mod, two(FICK, 1.0, 1.0, decay=0.3))
    assert np.allclose(v[:, 0], closed_two(FICK, 1.0, 1.0, decay=0.3), rtol=2e-6)


def test_fick_conserves_and_flattens(m, oracle_mod):
    rng = np.random.default_rng(4)
    N = 400
    x = rng.random((N, 3)) * 6.0
    w = resting(x, np.zeros(N, np.int32), ntypes=1)
    w["fluxes"] = [(FICK, 0, 0, 0, 0, 2.0, 0.0, 0.0)]
    s0 = rng.random(N).astype(np.float32) * 4
    w["species"] = np.stack([s0, np.zeros(N), np.zeros(N), np.zeros(N)], 1)
    o = oracle_mod.from_workload(w, m)
    assert o.step(50) == 0
    v, _ = o.species()
    assert abs(v[:, 0].sum(dtype=np.float64) - s0.sum(dtype=np.float64)) <= 1e-4 * s0.sum()
    assert v[:, 0].var() < 0.8 * s0.var()


def test_constant_species(m, oracle_mod):
    w = two(FICK, 2.0, 0.5)
    w["species_const"] = {1: 1}
    v, _ = run(m, oracle_mod, w, steps=3)
    assert v[1, 0] == np.float32(0.5)
    assert v[0, 0] < 2.0


def test_mover_keeps_flux(m, oracle_mod):
    """Particle 0 crosses x = 3 in its first step: its value is unchanged
    after that step, its accumulator holds that step's flux, and the next
    step integrates both steps' flux."""
    w = two(FICK, 2.0, 0.5)
    w["x"] = np.array([[2.999, 2.4, 2.5], [2.6, 2.6, 2.45]])
    w["v"] = np.array([[1.0, 0, 0], [0, 0, 0]])
    o = oracle_mod.from_workload(w, m)
    assert o.step(1) == 0
    v, q = o.species()
    assert v[0, 0] == np.float32(2.0) and q[0, 0] < 0
    assert v[1, 0] > 0.5 and q[1, 0] == 0
    q1 = q[0, 0]
    assert o.step(1) == 0
    v2, q2 = o.species()
    assert q2[0, 0] == 0
    assert v2[0, 0] < 2.0 + 0.01 * q1 * 1.5  # two steps of outflow
