"""Single-body forces on the GPU (MxForce.cpp:330-385) against the oracle:
Berendsen, constant and friction forces on an LJ fluid.  Bars as the other
parity tests: forces within 1e-5 relative (floor 0.1 rms), and the advance
relay (the oracle advancing with the GPU's forces) bit for bit."""
import numpy as np
import pytest

from mechanica_amd import workloads as W

pytestmark = pytest.mark.gpu


def with_force(kind):
    w = W.argon_lj(N=3000, seed=71, min_dist=0.3)
    if kind == "berendsen":
        ke = 0.5 * 39.4 * (w["v"].astype(np.float32).astype(np.float64) ** 2).sum()
        w["single_body"] = {0: (2, (0.05,))}
        w["target_energy"] = {0: 0.7 * ke}
    elif kind == "constant":
        w["single_body"] = {0: (3, np.array([0.5, -0.25, 1.0], np.float32))}
    else:
        w["single_body"] = {0: (4, (0.2, 0.5, 0.05, 3.0))}
    return w


@pytest.mark.parametrize("kind", ["berendsen", "constant", "friction"])
def test_single_body_parity(m, oracle_mod, kind):
    w = with_force(kind)
    print("case: single body", kind, flush=True)
    u = W.build_universe(m, w, seed=5)
    o = oracle_mod.from_workload(w, m, seed=5)
    u.step(1)
    o.force()
    fg = u.state()["f"].astype(np.float64)
    fr = o.state()["f"].astype(np.float64)
    n = np.linalg.norm(fr, axis=1)
    rel = (np.linalg.norm(fg - fr, axis=1) / np.maximum(n, 0.1 * np.sqrt((n ** 2).mean()))).max()
    print("rel(gpu,ref32)=%.3e" % rel)
    assert rel <= 1e-5
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


def test_berendsen_relaxes_to_target(m):
    """Free particles under the Berendsen force: the kinetic energy relaxes
    to the target (tau = 0.05, 300 steps of 0.01)."""
    rng = np.random.default_rng(8)
    N = 20000
    w = {"name": "free", "dim": [30.0] * 3, "cells": [10, 10, 10], "cutoff": 1.0, "dt": 0.01,
         "types": [{"mass": 2.0, "radius": 0.5, "dynamics": 0}],
         "x": rng.random((N, 3)) * 30.0, "v": rng.normal(0, 1, (N, 3)), "type": np.zeros(N, np.int32),
         "pots": [], "bonds": None}
    ke0 = 0.5 * 2.0 * (w["v"] ** 2).sum()
    w["single_body"] = {0: (2, (0.05,))}
    w["target_energy"] = {0: 2.0 * ke0}
    print("case: berendsen relaxation", flush=True)
    u = W.build_universe(m, w)
    u.step(300)
    ke = u.observables()["ekin"]
    assert abs(ke - 2.0 * ke0) <= 0.01 * 2.0 * ke0
