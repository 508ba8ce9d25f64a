"""Pins the oracle's Runge-Kutta 4 step (engine_advance_runge_kutta_4,
src/mdcore/src/engine_advance.cpp:443-698) against an independent FP64
Runge-Kutta integration of the same two-body ODE: a harmonic bond (whose
table is exact to float rounding) between two particles, Newtonian and
overdamped, including stages that cross cell boundaries (the shuffle moves
the particle and its start position p0 together)."""
import numpy as np
import pytest


def two_body(m, oracle_mod, x0, v0, dyn, dt=0.01, K=50.0, r0=1.0, integrator=1):
    w = {
        "name": "two_body", "dim": [6.0, 6.0, 6.0], "cells": [3, 3, 3], "cutoff": 1.0, "dt": dt,
        "types": [{"mass": 2.0, "radius": 0.5, "dynamics": dyn}],
        "x": np.asarray(x0, np.float64), "v": np.asarray(v0, np.float64), "type": np.zeros(2, np.int32),
        "pots": [],
        "bonds": {"i": np.array([0], np.int32), "j": np.array([1], np.int32),
                  "pot": ("harmonic", (0.0, 3.0, K, r0, 1e-4)), "dissociation": None},
        "integrator": integrator,
    }
    o = oracle_mod.from_workload(w, m)
    return o


def rk4_ref(x, v, dyn, dt, K, r0, mass, nsteps):
    """FP64 Runge-Kutta 4 of the two-body harmonic ODE in global coordinates."""
    x = np.array(x, np.float64)
    v = np.array(v, np.float64)

    def force(x):
        d = x[0] - x[1]
        r = np.linalg.norm(d)
        f0 = -2.0 * K * (r - r0) * d / r
        return np.stack([f0, -f0])

    for _ in range(nsteps):
        if dyn == 0:
            def deriv(x, v):
                return v, force(x) / mass
            k1x, k1v = deriv(x, v)
            k2x, k2v = deriv(x + 0.5 * dt * k1x, v + 0.5 * dt * k1v)
            k3x, k3v = deriv(x + 0.5 * dt * k2x, v + 0.5 * dt * k2v)
            k4x, k4v = deriv(x + dt * k3x, v + dt * k3v)
            x = x + dt / 6 * (k1x + 2 * k2x + 2 * k3x + k4x)
            v = v + dt / 6 * (k1v + 2 * k2v + 2 * k3v + k4v)
        else:
            k1 = force(x) / mass
            k2 = force(x + 0.5 * dt * k1) / mass
            k3 = force(x + 0.5 * dt * k2) / mass
            k4 = force(x + dt * k3) / mass
            x = x + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return x, v


@pytest.mark.parametrize("dyn", [0, 1])
@pytest.mark.parametrize("cross", [False, True])
def test_rk4_two_body(m, oracle_mod, dyn, cross):
    # cross: particle 0 sits just below a cell face (x = 2) and moves across
    # it within the step's stages.
    x0 = [[1.999 if cross else 1.5, 3.1, 3.2], [3.2, 3.0, 3.3]]
    v0 = [[0.4, 0.05, -0.02], [-0.3, 0.02, 0.01]] if dyn == 0 else [[0, 0, 0], [0, 0, 0]]
    dt, K, r0, mass = 0.01, 50.0, 1.0, 2.0
    for nsteps in (1, 25):
        o = two_body(m, oracle_mod, x0, v0, dyn, dt=dt, K=K, r0=r0)
        o.step(nsteps)
        assert o.error()[0] == 0
        xg = o.global_positions()
        xr, vr = rk4_ref(x0, v0, dyn, dt, K, r0, mass, nsteps)
        assert np.abs(xg - xr).max() < 2e-5 * max(1, nsteps / 5), (nsteps, np.abs(xg - xr).max())
        if dyn == 0:
            v = o.state()["v"].astype(np.float64)
            assert np.abs(v - vr).max() < 5e-5 * max(1, nsteps / 5)


def test_rk4_differs_from_euler(m, oracle_mod):
    """The integrator switch takes effect: one step of each differs by the
    Euler truncation error, and RK4 is the closer to the FP64 solution."""
    x0 = [[1.5, 3.1, 3.2], [3.2, 3.0, 3.3]]
    v0 = [[0.4, 0.05, -0.02], [-0.3, 0.02, 0.01]]
    xr, _ = rk4_ref(x0, v0, 0, 0.01, 50.0, 1.0, 2.0, 10)
    errs = []
    for integ in (0, 1):
        o = two_body(m, oracle_mod, x0, v0, 0, integrator=integ)
        o.step(10)
        errs.append(np.abs(o.global_positions() - xr).max())
    assert errs[1] < 1e-5 < errs[0]
