"""DPD pairs the engine evaluates but whose force is masked to zero: a DPD
cutoff below the engine's, the shifted DPD cutoff (b + R_i + R_j), and a
two-type universe mixing DPD (0-0, 1-1) with an LJ table (0-1), against the
oracle.  Bars as in test_gpu_parity: neighbour counts bit for bit; forces
within 1e-5 of the FP32 oracle per particle (floor 0.1 rms); the number
density within 1e-3 (the oracle keeps the reference's approximate rsqrt);
the virial within 1e-5 of the oracle's FP64 sum (largest entry); and a
five-step relay bit for bit."""
import numpy as np
import pytest

from mechanica_amd import workloads as W

pytestmark = pytest.mark.gpu


def case(name):
    # The reference's sigma term (deterministic): the oracle's FP64 virial
    # states it, not the counter noise.
    w = W.dpd_fluid(N=5000, sigma=3.0, seed=71)
    if name == "short_cutoff":
        w["pots"] = [("dpd", (25.0, 4.5, 3.0, 0.8, 0), 0, 0)]
    elif name == "shifted":
        w["types"] = [{"mass": 1.0, "radius": 0.2, "dynamics": 0}]
        w["pots"] = [("dpd", (25.0, 4.5, 3.0, 0.6, 1), 0, 0)]
    else:  # mixed DPD and LJ table
        rng = np.random.Generator(np.random.PCG64(72))
        w["types"] = [{"mass": 1.0, "radius": 0.1, "dynamics": 0}, {"mass": 1.0, "radius": 0.1, "dynamics": 0}]
        w["type"] = (rng.random(len(w["x"])) < 0.5).astype(np.int32)
        w["x"] = W._min_dist_fill(rng, len(w["x"]), w["dim"][0], 0.3)
        w["pots"] = [("dpd", (25.0, 4.5, 3.0, 1.0, 0), 0, 0), ("dpd", (25.0, 4.5, 3.0, 1.0, 0), 1, 1),
                     ("lj126", (0.275, 1.0, W.LJ_A, W.LJ_B, 1e-3), 0, 1)]
    w["name"] = "dpd_" + name
    return w


@pytest.mark.parametrize("name", ["short_cutoff", "shifted", "mixed"])
def test_dpd_cutoff_cases_match_oracle(m, oracle_mod, name):
    w = case(name)
    print("case: dpd cutoffs", name, flush=True)
    u = W.build_universe(m, w, seed=7)
    u.set_virial(True)
    o = oracle_mod.from_workload(w, m, seed=7)
    Wr = o.virial64()
    counts = u.neighbour_counts()
    u.step(1)
    g = u.state()
    o.force()
    r = o.state()
    assert np.array_equal(counts, r["ncount"])
    fr = r["f"].astype(np.float64)
    n = np.linalg.norm(fr, axis=1)
    floor = np.maximum(n, 0.1 * np.sqrt((n ** 2).mean()))
    rel = (np.linalg.norm(g["f"].astype(np.float64) - fr, axis=1) / floor).max()
    print("max rel force diff %.3e" % rel)
    assert rel <= 1e-5
    nd_g, nd_r = u.number_density(), o.number_density()
    assert np.abs(nd_g - nd_r).max() <= 1e-3 * np.abs(nd_r).max()
    Wg = u.observables()["virial"]
    assert np.abs(Wg - Wr).max() <= 1e-5 * np.abs(Wr).max()
    u2 = W.build_universe(m, w, seed=7)
    o2 = oracle_mod.from_workload(w, m, seed=7)
    for k in range(5):
        u2.step(1)
        s = u2.state()
        o2.step_with_forces(s["f"])
        t = o2.state()
        assert np.array_equal(s["cell"], t["cell"]), k
        assert np.array_equal(s["x"].view(np.uint32), t["x"].view(np.uint32)), k
        assert np.array_equal(s["v"].view(np.uint32), t["v"].view(np.uint32)), k
