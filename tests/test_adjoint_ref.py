"""Pins the FP64 adjoint reference (oracle/adjoint_ref.py) on the CPU.

* Morse closed form: minimum -D at s = 1, i.e. at r = R_i + R_j - r0;
* force = -dE/dx by FP64 central differences;
* trajectory gradients from autograd = FP64 central differences of the
  loss along random directions in x0 and theta, with the clamp inactive and
  with it active;
* a two-particle step computed by hand, clamp included.
"""
import math

import numpy as np
import pytest
import torch

from oracle.adjoint_ref import RefSystem
from mechanica_amd import workloads as W


def small_system(N=160, L=8.0, seed=3, min_dist=0.9):
    w = W.adjoint_morse(N=N, L=L, seed=seed, min_dist=min_dist)
    s = RefSystem(w["type"], w["dim"], w["cutoff"], w["dt"], mass=w["mass"], radius=w["radius"], smin=w["smin"],
                  smax=w["smax"])
    return s, w


def test_morse_minimum():
    s = RefSystem([0, 0], [10, 10, 10], 2.5, 0.01, radius=(1.0,))
    theta = torch.tensor([[2.0, 6.0, 1.3]], dtype=torch.float64)
    rmin = 2.0 - 1.3
    for r, want in ((rmin, -2.0),):
        x = torch.tensor([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]], dtype=torch.float64)
        assert abs(float(s.energy(x, theta)) - want) < 1e-12
        assert float(s.forces(x, theta).abs().max()) < 1e-12
    # Repulsive inside the minimum, attractive outside.
    x = torch.tensor([[1.0, 1.0, 1.0], [1.0 + rmin - 0.1, 1.0, 1.0]], dtype=torch.float64)
    assert float(s.forces(x, theta)[1, 0]) > 0
    x = torch.tensor([[1.0, 1.0, 1.0], [1.0 + rmin + 0.1, 1.0, 1.0]], dtype=torch.float64)
    assert float(s.forces(x, theta)[1, 0]) < 0


def test_force_is_minus_energy_gradient():
    s, w = small_system()
    x = torch.as_tensor(w["x"])
    theta = torch.as_tensor(w["theta"])
    F = s.forces(x, theta)
    rng = np.random.default_rng(0)
    eps = 1e-6
    for i in rng.choice(s.N, 6, replace=False):
        for k in range(3):
            xp, xm = x.clone(), x.clone()
            xp[i, k] += eps
            xm[i, k] -= eps
            fd = -(float(s.energy(xp, theta)) - float(s.energy(xm, theta))) / (2 * eps)
            assert abs(fd - float(F[i, k])) <= 1e-6 * max(1.0, abs(fd)), (i, k, fd, float(F[i, k]))


# Central-difference step: the six-step map is stiff (the directional
# derivative grows ~300x over the steps), so its second-order term is
# already visible at 1e-6.
EPS = 1e-7


def two_sided(f, x0, v, eps):
    return (float(f(x0 + eps * v)) - float(f(x0 - eps * v))) / (2 * eps)


@pytest.mark.parametrize("min_dist,nsteps", [(0.97, 6), (0.6, 6)])
def test_trajectory_gradient_matches_differences(min_dist, nsteps):
    """min_dist 0.97 keeps forces below the clamp (|dt F| < 0.05 h); 0.6
    starts with overlapping pairs whose steps are clamped."""
    s, w = small_system(min_dist=min_dist)
    target = torch.as_tensor(W.rdf_target(w, nbins=32, rmax=2.5))
    x0 = torch.as_tensor(w["x"]).clone().requires_grad_(True)
    theta = torch.as_tensor(w["theta"]).clone().requires_grad_(True)
    loss = s.loss(x0, theta, target, nsteps, nbins=32, rmax=2.5)
    gx, gt = torch.autograd.grad(loss, (x0, theta))
    rng = np.random.default_rng(1)
    with torch.no_grad():
        x0d, thd = x0.detach(), theta.detach()
        for _ in range(3):
            v = torch.as_tensor(rng.normal(size=x0d.shape))
            fd = two_sided(lambda x: s.loss(x, thd, target, nsteps, 32, 2.5), x0d, v, EPS)
            an = float((gx * v).sum())
            assert abs(fd - an) <= 1e-4 * max(abs(an), 1e-8), (fd, an)
            vt = torch.as_tensor(rng.normal(size=thd.shape))
            fd = two_sided(lambda t: s.loss(x0d, t, target, nsteps, 32, 2.5), thd, vt, EPS)
            an = float((gt * vt).sum())
            assert abs(fd - an) <= 1e-4 * max(abs(an), 1e-8), (fd, an)


def test_two_particle_step_by_hand():
    s = RefSystem([0, 0], [9, 9, 9], 2.5, 0.01, radius=(1.0,))
    theta = torch.tensor([[1.0, 6.0, 1.0]], dtype=torch.float64)
    for r in (1.3, 0.7):
        x = torch.tensor([[2.0, 2.0, 2.0], [2.0 + r, 2.0, 2.0]], dtype=torch.float64)
        u = math.exp(-6.0 * (r - 1.0))
        f = -2 * 6.0 * 1.0 * (u - u * u)  # -E'(r): force on particle 1 along +x
        dx = 0.01 * f
        maxx = float(np.float32(3.0 * 0.05))
        dx = max(min(dx, maxx), -maxx)
        x1 = s.step(x, theta)
        assert abs(float(x1[1, 0]) - (2.0 + r + dx)) < 1e-14
        assert abs(float(x1[0, 0]) - (2.0 - dx)) < 1e-14


# ------------------------------------------------ LJ pairs and harmonic bonds

def mixed_system(N=160, L=8.0, seed=3, min_dist=0.9, bond_r=1.1):
    w = W.adjoint_mixed(N=N, L=L, seed=seed, min_dist=min_dist, bond_r=bond_r)
    s = RefSystem(w["type"], w["dim"], w["cutoff"], w["dt"], mass=w["mass"], radius=w["radius"], smin=w["smin"],
                  smax=w["smax"], kinds=w["kinds"], bonds=w["bonds"])
    return s, w


def test_lj_pair_closed_form():
    """An A-B pair takes A s^-12 - B s^-6 with s = max(r, smin) and no
    radius shift; the force is -dE/dr along the pair; none beyond smax."""
    A_, B_ = 0.987, 1.987
    s = RefSystem([0, 1], [10, 10, 10], 2.5, 0.01, mass=(1.0, 1.0), radius=(1.0, 1.0),
                  kinds={(0, 1): (1, 0.5, 2.5)})
    theta = torch.tensor([[3.0, 6.0, 1.0], [A_, B_, 0.0], [3.0, 6.0, 1.0]], dtype=torch.float64)
    for r in (0.7, 1.0, 1.6, 0.3):
        x = torch.tensor([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]], dtype=torch.float64)
        re = max(r, 0.5)
        e = A_ / re ** 12 - B_ / re ** 6
        dedr = -12 * A_ / re ** 13 + 6 * B_ / re ** 7
        assert abs(float(s.energy(x, theta)) - e) <= 1e-12 * max(1.0, abs(e))
        # force on particle 1 along +x is -E'(re) * r / max(r, eps) / r ... = -E'(re)
        assert abs(float(s.forces(x, theta)[1, 0]) + dedr) <= 1e-10 * max(1.0, abs(dedr))
    x = torch.tensor([[1.0, 1.0, 1.0], [3.6, 1.0, 1.0]], dtype=torch.float64)
    assert float(s.energy(x, theta)) == 0.0


def test_bond_closed_form_and_clamp():
    """K (rb - r0)^2 with rb = clamp(r, bmin, bmax), force -E'(rb) / rb
    times the pair vector (bond_eval's clamp of r^2)."""
    s = RefSystem([0, 0], [20, 20, 20], 2.5, 0.01, radius=(1.0,), bonds=([0], [1], 0.2, 2.5))
    theta = torch.tensor([[0.0, 6.0, 1.0], [10.0, 1.0, 0.0]], dtype=torch.float64)  # no pair energy (d = 0)
    for r in (0.9, 1.3, 0.1, 3.0):
        x = torch.tensor([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]], dtype=torch.float64)
        rb = min(max(r, 0.2), 2.5)
        assert abs(float(s.energy(x, theta)) - 10.0 * (rb - 1.0) ** 2) < 1e-12
        f1 = -(2 * 10.0 * (rb - 1.0)) / rb * r  # x-force on particle 1 (d = x0 - x1 = -r)
        assert abs(float(s.forces(x, theta)[1, 0]) - f1) < 1e-10


def test_mixed_force_is_minus_energy_gradient():
    s, w = mixed_system()
    x = torch.as_tensor(w["x"])
    theta = torch.as_tensor(w["theta"])
    F = s.forces(x, theta)
    rng = np.random.default_rng(2)
    eps = 1e-6
    for i in rng.choice(s.N, 8, replace=False):
        for k in range(3):
            xp, xm = x.clone(), x.clone()
            xp[i, k] += eps
            xm[i, k] -= eps
            fd = -(float(s.energy(xp, theta)) - float(s.energy(xm, theta))) / (2 * eps)
            assert abs(fd - float(F[i, k])) <= 1e-6 * max(1.0, abs(fd)), (i, k, fd, float(F[i, k]))


def test_mixed_trajectory_gradient_matches_differences():
    """Gradients in x0 and in every theta row (Morse d, a, r0; LJ A, B; bond
    K, r0) against FP64 central differences."""
    s, w = mixed_system(min_dist=0.95)
    s.dt = 0.002
    target = torch.as_tensor(W.rdf_target(w, nbins=32, rmax=2.5))
    x0 = torch.as_tensor(w["x"]).clone().requires_grad_(True)
    theta = torch.as_tensor(w["theta"]).clone().requires_grad_(True)
    nsteps = 6
    loss = s.loss(x0, theta, target, nsteps, nbins=32, rmax=2.5)
    gx, gt = torch.autograd.grad(loss, (x0, theta))
    assert float(gt[1, :2].abs().sum()) > 0 and float(gt[3, :2].abs().sum()) > 0
    rng = np.random.default_rng(4)
    with torch.no_grad():
        x0d, thd = x0.detach(), theta.detach()
        for _ in range(3):
            v = torch.as_tensor(rng.normal(size=x0d.shape))
            fd = two_sided(lambda x: s.loss(x, thd, target, nsteps, 32, 2.5), x0d, v, EPS)
            an = float((gx * v).sum())
            assert abs(fd - an) <= 1e-4 * max(abs(an), 1e-8), (fd, an)
        for row, cols in ((1, (0, 1)), (3, (0, 1)), (0, (0, 1, 2))):
            for c in cols:
                e = torch.zeros_like(thd)
                e[row, c] = 1.0
                fd = two_sided(lambda t: s.loss(x0d, t, target, nsteps, 32, 2.5), thd, e, EPS)
                an = float(gt[row, c])
                assert abs(fd - an) <= 1e-4 * max(abs(an), 1e-8), (row, c, fd, an)
