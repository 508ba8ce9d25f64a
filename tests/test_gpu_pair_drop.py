"""Cell pairs that runner_dopair drops whole (its outer-loop break on a
negative key distance, runner_dopair.cpp:164-166) on the GPU against the
oracle: a sweep of four-particle cases across the x, y and z faces of two
non-cubic grids, and fluids with particles snapped to within 5e-5 of cell
faces, in the LDS kernel and in the global-memory fallback.

Neighbour counts are exact; forces within 1e-5 relative (floor 0.1 rms),
as in the other parity tests.  The fluids are checked to drop pairs (the
oracle counts fewer than a brute-force minimum-image count).
"""
import numpy as np
import pytest

from mechanica_amd import workloads as W
from test_oracle_pair_drop import brute_counts, four, workload
import test_oracle_pair_drop as T

pytestmark = pytest.mark.gpu


def rel_err(fg, fr):
    n = np.linalg.norm(fr, axis=1)
    return (np.linalg.norm(fg - fr, axis=1) / np.maximum(n, 0.1 * np.sqrt((n ** 2).mean()))).max()


def compare(m, oracle_mod, w, pmax=None):
    u = W.build_universe(m, w, seed=3)
    if pmax is not None:
        u.set_brick(2, 2, 2, 0, pmax)
    o = oracle_mod.from_workload(w, m, seed=3)
    counts = u.neighbour_counts()
    u.step(1)
    o.force()
    r = o.state()
    assert np.array_equal(counts, r["ncount"])
    fg, fr = u.state()["f"].astype(np.float64), r["f"].astype(np.float64)
    assert rel_err(fg, fr) <= 1e-5
    return r["ncount"]


@pytest.mark.parametrize("dim", [[6.0, 5.4, 6.6], [6.0, 5.7, 6.3]])
def test_four_particle_sweep(m, oracle_mod, monkeypatch, dim):
    monkeypatch.setattr(T, "DIM", dim)
    print("case: pair drop sweep dim", dim, flush=True)
    dropped = 0
    for axis in range(3):
        for e1 in np.linspace(2e-6, 1.2e-4, 5):
            for e2 in np.linspace(2e-6, 1.2e-4, 5):
                x = four(e1, e2, axis)
                n = compare(m, oracle_mod, workload(x))
                dropped += int(list(n) != list(brute_counts(x)))
    print("dropped cases:", dropped)
    assert dropped > 0


def snapped_fluid(N, seed, frac, gap=5e-5):
    rng = np.random.default_rng(seed)
    dim, cells = np.asarray(T.DIM), np.asarray(T.CELLS)
    h = dim / cells
    x = rng.random((N, 3)) * dim
    snap = rng.random((N, 3)) < frac
    lo = np.floor(x / h) * h
    low = rng.random((N, 3)) < 0.5
    g = rng.random((N, 3)) * gap
    return np.where(snap, np.where(low, lo + g, lo + h - g), x)


@pytest.mark.parametrize("frac", [0.01, 0.03])
@pytest.mark.parametrize("pmax", [None, 64])
def test_snapped_fluid(m, oracle_mod, frac, pmax):
    print("case: pair drop fluid frac", frac, "pmax", pmax, flush=True)
    x = snapped_fluid(3000, 7, frac)
    n = compare(m, oracle_mod, workload(x), pmax)
    assert (n < brute_counts(x)).sum() > 0


@pytest.mark.parametrize("pmax", [None, 64])
def test_window_rejections(m, oracle_mod, monkeypatch, pmax):
    """A random fluid on cells 1.0 x 0.9 x 1.1 with cutoff 1: runner_dopair's
    sort-key window rejects some pairs inside the cutoff (edge and corner
    neighbours of non-cubic cells); the pair phase tests the window on those
    offsets."""
    monkeypatch.setattr(T, "CUT", 1.0)
    print("case: window rejections pmax", pmax, flush=True)
    x = T.random_fluid(T.DIM)
    w = workload(x)
    w["pots"] = [("dpd", (25.0, 4.5, 0.0, 1.0, 0), 0, 0)]
    n = compare(m, oracle_mod, w, pmax)
    assert (n < brute_counts(x)).sum() > 0
