"""Runge-Kutta 4 step on the GPU (engine_advance_runge_kutta_4,
src/mdcore/src/engine_advance.cpp:443-698) against the oracle's RK4.

Each stage's forces differ from the oracle's in the last bits (summation
order), so states are compared with tolerances rather than bit for bit:
global positions within 2e-5 after one step and 2e-4 after five, velocities
within 1e-4 relative to the largest speed.  The two-body harmonic case is
also checked against an FP64 integration of its ODE, and graph replay of RK4
steps against eager launches bit for bit.
"""
import numpy as np
import pytest

from mechanica_amd import workloads as W

pytestmark = pytest.mark.gpu


def min_image(d, L):
    return d - L * np.round(d / L)


def rk4_cases():
    # The lattice at its Euler step (dt 0.005) is too stiff for the
    # unclamped RK4 step: the oracle itself stops with toofast (-29) in the
    # first step.  dt 0.002 is stable.
    lat = W.lattice(n=10, seed=62)
    lat["dt"] = 0.002
    return [
        W.argon_lj(N=3000, seed=61, min_dist=0.3),
        lat,
        W.dpd_fluid(N=4000, sigma=0.0, seed=63),
    ]


@pytest.mark.parametrize("wi", [0, 1, 2])
def test_rk4_matches_oracle(m, oracle_mod, wi):
    w = rk4_cases()[wi]
    w["integrator"] = 1
    print("case: rk4", w["name"], "N=%d" % len(w["x"]), flush=True)
    L = np.asarray(w["dim"])
    for nsteps, tol in ((1, 2e-5), (5, 2e-4)):
        u = W.build_universe(m, w, seed=3)
        o = oracle_mod.from_workload(w, m, seed=3)
        u.step(nsteps)
        o.step(nsteps)
        assert o.error()[0] == 0 and u.error == 0
        assert u.time == nsteps
        xg, vg = u.unload()
        xr = o.global_positions()
        vr = o.state()["v"].astype(np.float64)
        dx = np.abs(min_image(xg - xr, L)).max()
        dv = np.abs(vg - vr).max() / max(np.abs(vr).max(), 1e-30)
        print("steps %d  max|dx| %.3e  max|dv|/max|v| %.3e" % (nsteps, dx, dv))
        assert dx <= tol
        if w["types"][0]["dynamics"] == 0:
            assert dv <= 1e-4 * nsteps


def test_rk4_two_body_fp64(m):
    """Harmonic pair: the GPU's RK4 against an FP64 RK4 of the same ODE."""
    from test_oracle_rk4 import rk4_ref
    x0 = [[1.999, 3.1, 3.2], [3.2, 3.0, 3.3]]
    v0 = [[0.4, 0.05, -0.02], [-0.3, 0.02, 0.01]]
    w = {
        "name": "two_body", "dim": [6.0, 6.0, 6.0], "cells": [3, 3, 3], "cutoff": 1.0, "dt": 0.01,
        "types": [{"mass": 2.0, "radius": 0.5, "dynamics": 0}],
        "x": np.asarray(x0, np.float64), "v": np.asarray(v0, np.float64), "type": np.zeros(2, np.int32),
        "pots": [], "integrator": 1,
        "bonds": {"i": np.array([0], np.int32), "j": np.array([1], np.int32),
                  "pot": ("harmonic", (0.0, 3.0, 50.0, 1.0, 1e-4)), "dissociation": 1e30},
    }
    print("case: rk4 two body", flush=True)
    u = W.build_universe(m, w)
    u.step(25)
    xg, vg = u.unload()
    xr, vr = rk4_ref(x0, v0, 0, 0.01, 50.0, 1.0, 2.0, 25)
    assert np.abs(xg - xr).max() < 1e-4
    assert np.abs(vg - vr).max() < 2.5e-4


def test_rk4_toofast_like_oracle(m, oracle_mod):
    """The stiff lattice at dt 0.005: both stop with toofast in step one."""
    w = W.lattice(n=10, seed=62)
    w["integrator"] = 1
    print("case: rk4 lattice toofast", flush=True)
    u = W.build_universe(m, w, seed=3)
    o = oracle_mod.from_workload(w, m, seed=3)
    assert o.step(1) == -29
    with pytest.raises(m.EngineError) as e:
        u.step(1)
    assert e.value.code == -29
    assert u.error == -29


def test_rk4_graph_replay_bitwise(m):
    w = W.dpd_fluid(N=5000, sigma=3.0, seed=64)
    w["integrator"] = 1
    w["dpd_noise"] = True
    print("case: rk4 graphs", flush=True)
    out = []
    for graphs in (False, True):
        u = W.build_universe(m, w, seed=9)
        u.set_graphs(graphs)
        u.step(3)
        u.state()
        u.step(4)
        s = u.state()
        out.append((s["x"].copy(), s["v"].copy(), s["cell"].copy()))
    for a, b in zip(*out):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
