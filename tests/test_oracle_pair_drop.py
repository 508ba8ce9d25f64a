"""Pins the oracle's runner_dopair outer-loop break (runner_dopair.cpp:164-166).

The reference leaves its loop over the i cell when
(smallest j key) + dnshift - (i key) exceeds dmaxdist as an unsigned value.
A negative distance wraps to a huge one, so when the i particle with the
largest sort key already sees a negative distance, the whole cell pair
yields no pair, in the cutoff or not.  With key rounding this needs two
particles within about 1e-4 of the shared face and a cell width whose
discretised shift norm has a large fractional part (here the x face of
cells 1.0 x 0.9 x 1.1).

Checks against a brute-force minimum-image count: the dropped
configuration loses every cross-cell pair of that cell pair, the same
configuration with wider gaps or one particle moved away keeps all of them.
A random fluid on the same grid shows the window itself rejecting a few
pairs inside the cutoff, which a cubic grid never does.
"""
import sys

import numpy as np

DIM = [6.0, 5.4, 6.6]
CELLS = [6, 6, 6]
CUT = 0.85


def workload(x):
    x = np.asarray(x, np.float64)
    N = len(x)
    return {"name": "pair_drop", "dim": list(DIM), "cells": CELLS, "cutoff": CUT, "dt": 0.001,
            "types": [{"mass": 1.0, "radius": 0.1, "dynamics": 0}],
            "x": x, "v": np.zeros((N, 3)), "type": np.zeros(N, np.int32),
            "pots": [("dpd", (25.0, 4.5, 0.0, CUT, 0), 0, 0)], "bonds": None}


def four(e1, e2, axis=0, far=False):
    """a, c in one cell, b, d in the next one along axis; a and b at gaps
    e1, e2 from the shared face (b at 0.4 h from it when far).  b sits a
    quarter cell to the side of a, so no pair comes closer than the DPD
    potential's lower bound (sqrt(eps) = 3.5e-4, where the direction is
    random); the face's sort keys depend on the coordinate along axis only."""
    h = [DIM[k] / CELLS[k] for k in range(3)]
    face = 3 * h[axis]
    a = [2.5 * h[k] for k in range(3)]
    b, c, d = list(a), list(a), list(a)
    a[axis] = face - e1
    b[axis] = face + (0.4 * h[axis] if far else e2)
    c[axis] = face - 0.3 * h[axis]
    d[axis] = face + 0.3 * h[axis]
    b[(axis + 1) % 3] += 0.25 * h[(axis + 1) % 3]
    c[(axis + 1) % 3] += 0.2 * h[(axis + 1) % 3]
    d[(axis + 2) % 3] -= 0.1 * h[(axis + 2) % 3]
    return [a, b, c, d]


def brute_counts(x):
    x = np.asarray(x, np.float64)
    L = np.asarray(DIM)
    d = x[:, None, :] - x[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    return (r2 <= CUT * CUT).sum(1)


def counts(m, oracle_mod, x):
    o = oracle_mod.from_workload(workload(x), m)
    o.force()
    return np.asarray(o.state()["ncount"])


def test_cell_pair_dropped(m, oracle_mod):
    x = four(2e-6, 2e-6)
    assert list(brute_counts(x)) == [3, 3, 3, 3]
    # only the pairs inside each cell (a-c, b-d) survive
    assert list(counts(m, oracle_mod, x)) == [1, 1, 1, 1]


def test_drop_needs_both_particles_near_the_face(m, oracle_mod):
    for x in (four(3e-4, 3e-4), four(2e-6, 3e-4), four(3e-4, 2e-6), four(2e-6, 0, far=True)):
        assert list(counts(m, oracle_mod, x)) == list(brute_counts(x))


def test_drop_only_on_axes_whose_shift_rounds_up(m, oracle_mod):
    """The z face (cell width 1.1) has a discretised shift norm with a
    fractional part below one half: no gap drops its pairs."""
    for e in (2e-6, 1e-5, 4e-5):
        x = four(e, e, axis=2)
        assert list(counts(m, oracle_mod, x)) == list(brute_counts(x))


def random_fluid(dim, N=3000, seed=5):
    return np.random.default_rng(seed).random((N, 3)) * np.asarray(dim)


def test_window_rejects_in_cutoff_pairs_only_on_non_cubic_cells(m, oracle_mod, monkeypatch):
    """runner_dopair's window (key distance below dmaxdist) can reject a pair
    inside the cutoff only when |shift| exceeds shift.n, i.e. for edge and
    corner neighbours of non-cubic cells: on a cubic grid every in-cutoff
    pair's key distance stays below dscale * cutoff + 1 < dmaxdist."""
    here = sys.modules[__name__]
    monkeypatch.setattr(here, "CUT", 1.0)
    x = random_fluid(DIM)
    w = workload(x)
    w["pots"] = [("dpd", (25.0, 4.5, 0.0, 1.0, 0), 0, 0)]
    n = counts_of(m, oracle_mod, w)
    assert (n < brute_counts(x)).sum() > 0 and (n > brute_counts(x)).sum() == 0
    monkeypatch.setattr(here, "DIM", [6.0, 6.0, 6.0])
    x = random_fluid(here.DIM)
    w = workload(x)
    w["pots"] = [("dpd", (25.0, 4.5, 0.0, 1.0, 0), 0, 0)]
    assert list(counts_of(m, oracle_mod, w)) == list(brute_counts(x))


def counts_of(m, oracle_mod, w):
    o = oracle_mod.from_workload(w, m)
    o.force()
    return np.asarray(o.state()["ncount"])
