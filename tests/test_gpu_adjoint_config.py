"""The adjoint at the benchmarked configuration: 65,536 two-type Morse
particles in a periodic box of side 40, 50 overdamped steps at dt = 0.01,
checkpoints every 10 steps, loss the mean-squared deviation of the final
64-bin g(r) from a seeded target.

At dt = 0.01 the map is unstable near contact (tests/test_gpu_adjoint.py
compares with the FP64 reference at dt = 0.002 for that reason), so this
size is checked through properties that hold for any trajectory:
  * translation invariance: the loss depends on pair distances only and
    every step's Jacobian maps a uniform shift to itself (Newton's third
    law), so sum_i dL/dx0_i = 0 for each component.  A dropped, doubled or
    wrong-signed pair term breaks it at the size of the gradient itself;
    the bar is 1e-6 of sum_i |dL/dx0_i|;
  * the gradients are the same bit for bit with checkpoints every 10 and
    every 25 steps (the recomputation repeats the forward exactly);
  * the loss from the differentiable run equals the forward-only run's.
"""
import numpy as np
import pytest
import torch

from mechanica_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_adjoint_benchmark_config_properties():
    import mechanica_amd.adjoint as A
    w = W.adjoint_morse()
    print("case: adjoint N=%d dt=%g T=50 bins=64" % (len(w["x"]), w["dt"]), flush=True)
    sysm = A.MorseSystem(w["type"], w["dim"], w["cutoff"], w["dt"], mass=w["mass"], radius=w["radius"],
                         smin=w["smin"], smax=w["smax"])
    x32 = torch.as_tensor(w["x"].astype(np.float32).astype(np.float64))
    theta = torch.as_tensor(w["theta"])
    target = torch.as_tensor(np.random.default_rng(21).uniform(0.5, 1.5, size=(3, 64))).cuda()
    grads, losses = [], []
    for every in (10, 25):
        x0 = x32.clone().cuda().requires_grad_(True)
        th = theta.clone().requires_grad_(True)
        loss = A.rdf_loss(sysm, x0, th, target, 50, every=every, nbins=64, rmax=2.5)
        loss.backward()
        grads.append((x0.grad.detach().cpu().numpy(), th.grad.detach().cpu().numpy()))
        losses.append(float(loss.detach()))
    gx, gt = grads[0]
    assert np.isfinite(gx).all() and np.isfinite(gt).all()
    drift = np.abs(gx.sum(0))
    scale = np.abs(gx).sum(0)
    print("loss %.12g  sum dL/dx0 %s  sum |dL/dx0| %s" % (losses[0], drift, scale))
    assert (scale > 0).all()
    assert (drift <= 1e-6 * scale).all()
    assert np.array_equal(grads[1][0], gx) and np.array_equal(grads[1][1], gt)
    assert losses[1] == losses[0]
    with torch.no_grad():
        xT = A.trajectory(sysm, x32.clone().cuda(), theta, 50, 10)
        g = A.rdf_normalise(A.soft_rdf(sysm, xT, 64, 2.5), sysm, 2.5)
        plain = float(((g - target) ** 2).mean())
    assert plain == losses[0]
