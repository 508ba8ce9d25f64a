"""Species transport on the GPU (flux_eval_ex, flux_eval.hpp:33-81;
MxFluxes_Integrate, Flux.cpp:261-284) against the oracle on a two-type DPD
fluid with Fick, secrete and uptake fluxes and decay.

The per-particle flux sums run in a different order than the oracle's, so
species values are compared within 1e-5 of the largest value; positions,
velocities and cells (the oracle relaying the GPU's forces) stay bit for
bit, which also checks that movers keep their flux for the next step.
"""
import numpy as np
import pytest

from mechanica_amd import workloads as W

pytestmark = pytest.mark.gpu

FICK, SECRETE, UPTAKE = 0, 1, 2


def species_fluid(N=6000, seed=31, fluxes=True):
    rng = np.random.default_rng(seed)
    w = W.dpd_fluid(N=N, sigma=0.0, seed=seed)
    w["types"] = [dict(w["types"][0]), dict(w["types"][0])]
    w["type"] = (rng.random(N) < 0.4).astype(np.int32)
    w["pots"] = [("dpd", (25.0, 4.5, 0.0, 1.0, 0), 0, 0), ("dpd", (25.0, 4.5, 0.0, 1.0, 0), 0, 1),
                 ("dpd", (25.0, 4.5, 0.0, 1.0, 0), 1, 1)]
    w["species"] = rng.random((N, 3)).astype(np.float32) * [2.0, 1.0, 3.0]
    if fluxes:
        w["fluxes"] = [(FICK, 0, 0, 0, 0, 1.0, 0.0, 0.0),
                       (FICK, 0, 1, 0, 0, 0.5, 0.1, 0.0),
                       (SECRETE, 0, 1, 1, 1, 0.3, 0.0, 0.5),
                       (UPTAKE, 1, 0, 2, 2, 0.2, 0.05, 2.0)]
        w["species_const"] = {1: 0b100}
    return w


def close(a, b, tol=1e-5):
    scale = max(np.abs(b).max(), 1.0)
    return np.abs(a - b).max() <= tol * scale


def test_fluxes_relay(m, oracle_mod):
    w = species_fluid()
    print("case: fluxes relay N=%d" % len(w["x"]), flush=True)
    u = W.build_universe(m, w, seed=5)
    o = oracle_mod.from_workload(w, m, seed=5)
    for k in range(20):
        u.step(1)
        g = u.state()
        o.step_with_forces(g["f"])
        r = o.state()
        assert np.array_equal(g["cell"], r["cell"]), k
        assert np.array_equal(g["x"].view(np.uint32), r["x"].view(np.uint32)), k
        vg, qg = u.species()
        vr, qr = o.species()
        assert close(vg, vr), (k, np.abs(vg - vr).max())
        assert close(qg, qr), (k, np.abs(qg - qr).max())
    # the constant species of type 1 never changed
    s0 = np.asarray(w["species"], np.float32)
    t1 = w["type"] == 1
    assert np.array_equal(vg[t1, 2], s0[t1, 2])


def test_species_without_fluxes_are_carried(m):
    w = species_fluid(fluxes=False)
    print("case: species carried", flush=True)
    u = W.build_universe(m, w, seed=5)
    u.step(30)
    v, q = u.species()
    assert np.array_equal(v[:, :3], np.asarray(w["species"], np.float32))
    assert np.all(q == 0)


def test_fluxes_graph_replay_bitwise(m):
    w = species_fluid(N=4000, seed=32)
    print("case: fluxes graphs", flush=True)
    out = []
    for graphs in (False, True):
        u = W.build_universe(m, w, seed=9)
        u.set_graphs(graphs)
        u.step(3)
        u.state()
        u.step(4)
        out.append(u.species())
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_fluxes_rk4_accumulate(m, oracle_mod):
    """Under RK4 the reference never integrates: values stay, the flux of
    the four stages accumulates."""
    w = species_fluid(N=3000, seed=33)
    w["integrator"] = 1
    print("case: fluxes rk4", flush=True)
    u = W.build_universe(m, w, seed=5)
    o = oracle_mod.from_workload(w, m, seed=5)
    u.step(1)
    assert o.step(1) == 0
    vg, qg = u.species()
    vr, qr = o.species()
    assert np.array_equal(vg[:, :3], np.asarray(w["species"], np.float32))
    assert close(qg, qr, 1e-4)
