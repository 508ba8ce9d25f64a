"""The oracle (CPU restatement of mdcore's step) pinned without the reference.

No test, fixture or golden vector in the reference touches this path, and
the reference cannot be compiled here, so the oracle is checked against what
the mathematics fixes, chosen so plausible mistakes fail:
  * neighbour counts against brute-force O(N^2) enumeration with the
    minimum image (a wrong shift sign, a missed cell pair or the pruning
    window cutting real pairs changes a count);
  * two-particle forces against the closed-form LJ / Morse / DPD / harmonic
    formulas, across a cell face and across the periodic seam (a sign or
    orientation error flips the force);
  * Newton's third law (sum of pair forces is zero) and momentum
    conservation of DPD at sigma = 0;
  * single-particle advance cases computed by hand: migration to the next
    cell, the 5% velocity clamp, the overdamped displacement clamp, frozen
    axes, and the too-fast error;
  * bond dissociation at the threshold (force / r compared with the energy,
    bond.cpp:227-232);
  * the threaded runner agrees with the serial one.
"""
import math

import numpy as np
import pytest

from mechanica_amd import workloads as W


def brute_counts(x, L, rc, margin):
    D = x[:, None, :] - x[None, :, :]
    D -= L * np.round(D / L)
    r2 = (D ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    lo = (r2 <= rc * rc - margin).sum(1)
    hi = (r2 <= rc * rc + margin).sum(1)
    return lo, hi


@pytest.mark.parametrize("which", ["argon", "dpd", "lattice"])
def test_neighbour_counts_brute_force(m, oracle_mod, which):
    if which == "argon":
        w = W.argon_lj(N=1500, seed=31, min_dist=0.275)
    elif which == "dpd":
        w = W.dpd_fluid(N=1500, seed=32)
    else:
        w = W.lattice(n=9, seed=33)
    o = oracle_mod.from_workload(w, m)
    o.force()
    nc = o.state()["ncount"]
    lo, hi = brute_counts(o.global_positions(), w["dim"][0], w["cutoff"], 1e-5)
    assert np.all(nc >= lo) and np.all(nc <= hi)
    assert nc.sum() > 0


def two_body(m, oracle_mod, pot, x1, x2, v1=(0, 0, 0), v2=(0, 0, 0), L=6.0, cells=6, rc=1.0, mass=1.0,
             radius=1.0, dt=0.01):
    o = oracle_mod.Oracle([L] * 3, [cells] * 3, rc, dt)
    o.add_type(mass, radius, 0)
    o.bind(pot, 0, 0)
    o.load(np.array([x1, x2], float), np.array([v1, v2], float), np.zeros(2, np.int32))
    o.force()
    return o.state()["f"].astype(np.float64), o


def lj_force(r):
    return -(-12 * W.LJ_A / r ** 13 + 6 * W.LJ_B / r ** 7)  # -dE/dr


@pytest.mark.parametrize("r", [0.3, 0.45, 0.7, 0.95])
@pytest.mark.parametrize("where", ["inside", "face", "seam"])
def test_two_body_lj(m, oracle_mod, r, where):
    pot = m.Potential.lennard_jones_12_6(0.275, 1.0, W.LJ_A, W.LJ_B, 1e-3)
    base = {"inside": 2.1, "face": 2.9 - r / 2 + 0.05, "seam": 6.0 - r / 2}[where]
    x1 = np.array([base, 3.3, 3.4])
    x2 = x1 + np.array([r, 0.0, 0.0])
    x2[0] %= 6.0
    f, _ = two_body(m, oracle_mod, pot, x1, x2, radius=0.1)
    fx = lj_force(r)  # force on particle 2 (+x side) along +x
    assert f[1, 0] == pytest.approx(fx, rel=0.1, abs=0.05)
    assert f[0, 0] == pytest.approx(-f[1, 0], rel=1e-6)
    assert abs(f[:, 1:]).max() <= 1e-6 * abs(f[1, 0]) + 1e-9


@pytest.mark.parametrize("r", [0.9, 1.0, 1.2, 1.6])
def test_two_body_morse_shift(m, oracle_mod, r):
    pot = m.Potential.morse(d=1.0, a=6.0, r0=1.0, min=1e-4, max=2.0, tol=1e-3)
    x1 = np.array([1.3, 2.2, 3.1])
    d = np.array([0.6, -0.48, 0.64])  # unit vector
    x2 = x1 + r * d
    f, _ = two_body(m, oracle_mod, pot, x1, x2, L=8.0, cells=4, rc=2.0)
    # radius 1 each, r0 = 1: the shifted argument equals r, minimum at r = 1
    dE = 1.0 * (-12 * np.exp(-12 * (r - 1)) + 12 * np.exp(-6 * (r - 1)))
    f2 = -dE * d
    assert np.allclose(f[1], f2, rtol=0.1, atol=0.05 + 0.02 * abs(dE))
    assert np.allclose(f[0], -f[1], rtol=1e-6, atol=1e-7)


def test_two_body_dpd(m, oracle_mod):
    alpha, gamma, sigma, dt = 25.0, 4.5, 3.0, 0.01
    pot = m.Potential.dpd(alpha=alpha, gamma=gamma, sigma=sigma, cutoff=1.0)
    x1 = np.array([2.2, 2.5, 2.5])
    x2 = np.array([2.8, 2.9, 2.5])
    v1 = np.array([0.3, -0.2, 0.1])
    v2 = np.array([-0.1, 0.4, 0.0])
    f, _ = two_body(m, oracle_mod, pot, x1, x2, v1, v2, L=5.0, cells=5, dt=dt)
    dx = x1 - x2
    r = np.linalg.norm(dx)
    e = dx / r
    w = 1 - r
    fs = alpha * w - gamma * w * w * (e @ (v1 - v2)) + sigma * w / math.sqrt(dt)
    assert np.allclose(f[0], fs * e, rtol=1e-5)
    assert np.allclose(f[1], -fs * e, rtol=1e-5)


def test_pair_forces_sum_to_zero(m, oracle_mod):
    w = W.argon_lj(N=2000, seed=34, min_dist=0.275)
    o = oracle_mod.from_workload(w, m)
    o.force()
    f = o.state()["f"].astype(np.float64)
    assert np.abs(f.sum(0)).max() <= 1e-5 * np.abs(f).sum()


def test_dpd_momentum_conserved(m, oracle_mod):
    w = W.dpd_fluid(N=3000, sigma=0.0, seed=35)
    o = oracle_mod.from_workload(w, m)
    p0 = o.state()["v"].astype(np.float64).sum(0)
    o.step(10)
    p1 = o.state()["v"].astype(np.float64).sum(0)
    assert np.abs(p1 - p0).max() <= 1e-3


def one_particle(oracle_mod, x, v, dyn=0, flags=None, frac=None, h=1.0, dt=0.01):
    o = oracle_mod.Oracle([5 * h] * 3, [5] * 3, 1.0 * h, dt)
    o.add_type(2.0, 0.1, dyn)
    if frac is not None:
        o.set_max_dist_fraction(frac)
    o.load(np.array([x], float), np.array([v], float), np.zeros(1, np.int32),
           None if flags is None else np.array([flags], np.uint16))
    return o


def test_advance_migrates_to_next_cell(oracle_mod):
    o = one_particle(oracle_mod, [1.995, 2.5, 2.5], [1.0, 0.0, 0.0])
    o.step(1)
    s = o.state()
    x0 = np.float32(1.995) - np.float32(1.0)  # local coordinate in cell 1
    x1 = np.float32(x0 + np.float32(0.01) * np.float32(1.0)) - np.float32(1.0)
    assert s["cell"][0] == (2 * 5 + 2) * 5 + 2
    assert s["x"][0, 0] == x1


def test_advance_velocity_clamp(oracle_mod):
    o = one_particle(oracle_mod, [2.5, 2.5, 2.5], [20.0, -30.0, 1.0])
    o.step(1)
    v = o.state()["v"][0]
    maxv = np.float32(np.float32(1.0) * np.float32(0.05)) / np.float32(0.01)
    assert v[0] == maxv and v[1] == -maxv and v[2] == np.float32(1.0)


def test_advance_overdamped_clamp_and_frozen(m, oracle_mod):
    pot = m.Potential.lennard_jones_12_6(0.275, 1.0, W.LJ_A, W.LJ_B, 1e-3)
    o = oracle_mod.Oracle([5.0] * 3, [5] * 3, 1.0, 0.01)
    o.add_type(1.0, 0.1, 1)
    o.bind(pot, 0, 0)
    x = np.array([[2.5, 2.5, 2.5], [2.8, 2.5, 2.5]])
    o.load(x, np.zeros_like(x), np.zeros(2, np.int32), np.array([0, m.FROZEN_X], np.uint16))
    before = o.state()["x"].copy()
    o.step(1)
    s = o.state()
    # particle 0 is pushed to -x by the strong repulsion; the step is clamped
    # to 5% of the cell width
    assert s["x"][0, 0] == np.float32(before[0, 0] - np.float32(0.05))
    # particle 1 is frozen in x
    assert s["x"][1, 0] == before[1, 0]


def test_particle_flag_values(m):
    """PARTICLE_FROZEN_X/Y/Z are 1 << 3, 1 << 4, 1 << 5 (MxParticle.h:58-60;
    1 << 2 is PARTICLE_BOUND), in the Python module and the C header."""
    assert (m.FROZEN_X, m.FROZEN_Y, m.FROZEN_Z) == (1 << 3, 1 << 4, 1 << 5)
    import os
    import re
    text = open(os.path.join(os.path.dirname(__file__), "..", "include", "mxg.h")).read()
    vals = {k: int(v) for k, v in re.findall(r"#define MXG_FROZEN_([XYZ]) \(1 << (\d+)\)", text)}
    assert vals == {"X": 3, "Y": 4, "Z": 5}


def test_advance_frozen_axes_and_bound_flag(oracle_mod):
    """A particle frozen in y keeps its y and moves in x and z; PARTICLE_BOUND
    (1 << 2) does not freeze anything."""
    for flags, still in ((1 << 4, [1]), (1 << 2, []), ((1 << 3) | (1 << 5), [0, 2])):
        o = one_particle(oracle_mod, [2.5, 2.5, 2.5], [1.0, -2.0, 3.0], flags=flags)
        before = o.state()["x"][0].copy()
        o.step(1)
        after = o.state()["x"][0]
        for k in range(3):
            assert (after[k] == before[k]) == (k in still), (flags, k)


def test_local_shuffle_at_step_start(oracle_mod):
    """space_shuffle_local (space.cpp:333-393) at the start of the step's
    force phase, by hand: x - origin truncates toward zero, so x = -1e-9 loads
    into cell 0 at -1e-9 and moves to the last cell at -1e-9 + h, which rounds
    to h; x = L - 1e-9 loads into the last cell at 1 - 1e-9, which rounds to
    h, and moves to cell 0 at h - h = 0."""
    o = oracle_mod.Oracle([5.0] * 3, [5] * 3, 1.0, 0.01)
    o.add_type(1.0, 0.1, 0)
    x = np.array([[-1e-9, 2.5, 2.5], [5.0 - 1e-9, 2.5, 2.5]])
    o.load(x, np.zeros_like(x), np.zeros(2, np.int32))
    s = o.state()
    assert s["cell"].tolist() == [(0 * 5 + 2) * 5 + 2, (4 * 5 + 2) * 5 + 2]
    assert s["x"][0, 0] == np.float32(-1e-9) and s["x"][1, 0] == np.float32(1.0)
    o.force()
    s = o.state()
    assert s["cell"].tolist() == [(4 * 5 + 2) * 5 + 2, (0 * 5 + 2) * 5 + 2]
    assert s["x"][0, 0] == np.float32(1.0) and s["x"][1, 0] == np.float32(0.0)


def test_toofast_error(oracle_mod):
    o = one_particle(oracle_mod, [2.5, 2.5, 2.5], [250.0, 0.0, 0.0], frac=3.0)
    rc = o.step(1)
    assert rc == -29
    assert o.error() == (-29, 0)
    assert o.step(1) == -29  # sticky: later steps are refused


@pytest.mark.parametrize("r,breaks", [(1.02, False), (1.03, True)])
def test_bond_dissociation(m, oracle_mod, r, breaks):
    hp = m.Potential.harmonic(100.0, 1.0)
    o = oracle_mod.Oracle([8.0] * 3, [4] * 3, 2.0, 0.005)
    o.add_type(1.0, 1.0, 1)
    x = np.array([[3.0, 3.5, 3.5], [3.0 + r, 3.5, 3.5]])
    o.load(x, np.zeros_like(x), np.zeros(2, np.int32))
    o.set_bonds([0], [1], hp, 5.0)
    o.force()
    f = o.state()["f"][:, 0]
    assert o.bond_active()[0] == (not breaks)
    if breaks:
        assert f[0] == 0 and f[1] == 0
    else:
        assert f[0] == pytest.approx(2 * 100 * (r - 1), rel=2e-3)
        assert f[1] == pytest.approx(-f[0], rel=1e-6)


def test_threaded_runner_matches_serial(m, oracle_mod):
    w = W.argon_lj(N=3000, seed=36, min_dist=0.275)
    a = oracle_mod.from_workload(w, m, threads=1)
    b = oracle_mod.from_workload(w, m, threads=4)
    a.force()
    b.force()
    fa, fb = a.state()["f"], b.state()["f"]
    n = np.linalg.norm(fa, axis=1)
    floor = np.maximum(n, 0.1 * np.sqrt((n ** 2).mean()))
    assert (np.linalg.norm(fa - fb, axis=1) / floor).max() <= 1e-5
    assert np.array_equal(a.state()["ncount"], b.state()["ncount"])


def test_philox_known_answers(oracle_mod):
    """Random123's published known-answer vectors for Philox4x32-10."""
    assert oracle_mod.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle_mod.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert oracle_mod.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


@pytest.mark.parametrize("r", [0.35, 0.7, 0.95, 1.05])
def test_two_body_dpd_shifted(m, oracle_mod, r):
    """Shifted DPD (dpd_eval.hpp:21-69): cutoff b + R_i + R_j, omega_c from
    r - R_i - R_j (1 inside the radii); with b = 0.6 and radii 0.2 the
    cutoff is 1.0; a pair beyond it gets no force."""
    alpha, gamma, sigma, dt, b, R = 25.0, 4.5, 3.0, 0.01, 0.6, 0.2
    pot = m.Potential.dpd(alpha=alpha, gamma=gamma, sigma=sigma, cutoff=b, shifted=True)
    d = np.array([0.6, -0.48, 0.64])
    x1 = np.array([2.2, 2.5, 2.5])
    x2 = x1 + r * d
    v1 = np.array([0.3, -0.2, 0.1])
    v2 = np.array([-0.1, 0.4, 0.0])
    f, _ = two_body(m, oracle_mod, pot, x1, x2, v1, v2, L=5.0, cells=5, rc=1.2, radius=R, dt=dt)
    cut = b + 2 * R
    if r > cut:
        assert np.abs(f).max() == 0.0
        return
    e = (x1 - x2) / r
    sr = r - 2 * R
    w = 1.0 if sr < 0 else 1 - sr / cut
    fs = alpha * w - gamma * w * w * (e @ (v1 - v2)) + sigma * w / math.sqrt(dt)
    assert np.allclose(f[0], fs * e, rtol=1e-5, atol=1e-6)
    assert np.allclose(f[1], -fs * e, rtol=1e-5, atol=1e-6)
