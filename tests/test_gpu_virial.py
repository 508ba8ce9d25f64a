"""The engine's virial tensor (mxg_engine_set_virial: sum over pairs and
bonds of (x_i - x_j) (x) f_i, accumulated in FP64) against the oracle's FP64
statement of the same sum, whose definition is pinned in
test_oracle_virial.py.  Bar: 1e-5 of the tensor's largest entry (the FP32
pair forces differ from FP64 at that level)."""
import numpy as np
import pytest

from mechanica_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["lj", "dpd", "lattice"])
def test_virial_matches_fp64(m, oracle_mod, name):
    if name == "lj":
        w = W.argon_lj(N=4000, seed=81, min_dist=0.3)
    elif name == "dpd":
        w = W.dpd_fluid(N=6000, sigma=0.0, seed=82)
    else:
        w = W.lattice(n=12, seed=83)
    print("case: virial", name, flush=True)
    u = W.build_universe(m, w, seed=5)
    u.set_virial(True)
    o = oracle_mod.from_workload(w, m, seed=5)
    Wr = o.virial64()
    u.step(1)
    Wg = u.observables()["virial"]
    print(Wg, Wr)
    assert np.abs(Wg - Wr).max() <= 1e-5 * np.abs(Wr).max()
