"""Pins the oracle's single-body forces (MxForce.cpp:330-385) on free
particles (no pair potentials), where one Euler step has a closed form:
  constant  v' = v + dt F / m (the advance's float/double mix, bit for bit);
  friction  v' = v + dt (-coef |v|^2 v) / m   (random part off: mean = std = 0);
  berendsen v' = v + dt itau (E_target / E_kin - 1) v / m, and over many
            steps E_kin relaxes to E_target.
"""
import numpy as np


def free_particles(N=200, seed=5, dim=8.0):
    rng = np.random.default_rng(seed)
    return {
        "name": "free", "dim": [dim] * 3, "cells": [3, 3, 3], "cutoff": 1.0, "dt": 0.01,
        "types": [{"mass": 2.0, "radius": 0.5, "dynamics": 0}],
        "x": rng.random((N, 3)) * dim, "v": rng.normal(0, 1, (N, 3)), "type": np.zeros(N, np.int32),
        "pots": [], "bonds": None,
    }


def euler_v(v, f, dt, imass):
    """The advance's velocity update: float(double(v) + double(float(dt * f)) * imass)."""
    v = v.astype(np.float32)
    f = f.astype(np.float32)
    dtf = (np.float32(dt) * f).astype(np.float32)
    return (v.astype(np.float64) + dtf.astype(np.float64) * imass).astype(np.float32)


def run(m, oracle_mod, w, steps=1):
    o = oracle_mod.from_workload(w, m)
    v0 = o.state()["v"].copy()
    o.step(steps)
    return o, v0


def test_constant_force_bitwise(m, oracle_mod):
    w = free_particles()
    F = np.array([0.3, -1.25, 2.0], np.float32)
    w["single_body"] = {0: (3, F)}
    o, v0 = run(m, oracle_mod, w)
    ref = euler_v(v0, np.broadcast_to(F, v0.shape), 0.01, 0.5)
    assert np.array_equal(o.state()["v"].view(np.uint32), ref.view(np.uint32))


def test_friction_closed_form(m, oracle_mod):
    w = free_particles()
    coef = 0.7
    w["single_body"] = {0: (4, (0.0, 0.0, 0.01, coef))}
    o, v0 = run(m, oracle_mod, w)
    v = v0.astype(np.float64)
    f = -coef * (v * v).sum(1, keepdims=True) * v
    ref = v + 0.01 * f * 0.5
    assert np.abs(o.state()["v"] - ref).max() <= 1e-6 * np.abs(ref).max()


def test_berendsen_closed_form_and_relaxation(m, oracle_mod):
    w = free_particles(N=400)
    tau = 0.05
    ke0 = 0.5 * 2.0 * (w["v"].astype(np.float32).astype(np.float64) ** 2).sum()
    target = 0.5 * ke0
    w["single_body"] = {0: (2, (tau,))}
    w["target_energy"] = {0: target}
    o, v0 = run(m, oracle_mod, w)
    v = v0.astype(np.float64)
    ke = (0.5 * 2.0 * (v0.astype(np.float32) ** 2).sum(1)).astype(np.float64).sum()
    scale = (1 / tau) * (target / ke - 1.0)
    ref = v + 0.01 * scale * v * 0.5
    assert np.abs(o.state()["v"] - ref).max() <= 1e-6 * np.abs(ref).max()
    o.step(300)
    assert abs(o.observables()["ekin"] - target) <= 0.01 * target
