"""mxg_morse_forces_noalloc (the entry point that allocates nothing) against
mxg_adjoint_forces on the same inputs, through the C ABI.

Both run the same binning and force kernels, so the forces agree bit for bit
and the FP64 energies (summed on the device in a fixed order against the
host's sum of the same partials) to 1e-12 relative.  Two calls in a row, the
second with fewer particles and a smaller grid, show that the scratch left by
the first call does not leak into the second; sizes beyond the scratch are
refused with MXG_ERR_RANGE.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from mechanica_amd import _capi
from mechanica_amd import adjoint as ADJ
from mechanica_amd import workloads as W

pytestmark = pytest.mark.gpu


def _d(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(_capi.c_double_p)


def noalloc_forces(w, cells, theta):
    L = _capi.lib()
    N = len(w["x"])
    dev = torch.device("cuda")
    x = torch.as_tensor(np.ascontiguousarray(w["x"], np.float64), device=dev)
    t = torch.as_tensor(np.ascontiguousarray(w["type"], np.int32), device=dev)
    f = torch.full((N, 3), float("nan"), dtype=torch.float32, device=dev)
    e = torch.full((1,), float("nan"), dtype=torch.float64, device=dev)
    dim = np.ascontiguousarray(w["dim"], np.float64)
    c = np.ascontiguousarray(cells, np.int32)
    rad = np.ascontiguousarray(w["radius"], np.float64)
    th = np.ascontiguousarray(theta, np.float64).reshape(-1)
    stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    rc = L.mxg_morse_forces_noalloc(N, dim.ctypes.data_as(_capi.c_double_p), c.ctypes.data_as(_capi.c_int_p),
                                    float(w["cutoff"]), len(w["radius"]), rad.ctypes.data_as(_capi.c_double_p),
                                    th.ctypes.data_as(_capi.c_double_p), float(w["smin"]), float(w["smax"]),
                                    C.c_void_p(t.data_ptr()), C.c_void_p(x.data_ptr()), C.c_void_p(f.data_ptr()),
                                    C.c_void_p(e.data_ptr()), stream)
    torch.cuda.synchronize()
    return rc, f.cpu().numpy(), float(e.cpu()[0])


def reference_forces(w, cells, theta):
    sysm = ADJ.MorseSystem(w["type"], w["dim"], w["cutoff"], w["dt"], mass=w["mass"], radius=w["radius"],
                           cells=cells, smin=w["smin"], smax=w["smax"])
    f, e = sysm.forces(torch.as_tensor(w["x"]), theta)
    torch.cuda.synchronize()
    return f.cpu().numpy(), e


def test_noalloc_matches_adjoint_forces_twice():
    big = W.adjoint_morse(N=60000, L=36.0, seed=61)
    small = W.adjoint_morse(N=7000, L=18.0, seed=62)
    theta = np.array([[3.0, 6.0, 1.0], [1.0, 6.0, 1.0], [3.0, 6.0, 1.0]])
    for w, cells in ((big, [14, 14, 14]), (small, [7, 7, 7])):
        print("case: no-alloc Morse forces N=%d cells=%s" % (len(w["x"]), cells), flush=True)
        rc, f, e = noalloc_forces(w, cells, theta)
        assert rc == 0
        fr, er = reference_forces(w, cells, theta)
        assert np.array_equal(f.view(np.uint32), fr.view(np.uint32))
        assert abs(e - er) <= 1e-12 * abs(er)
        print("energy %.12g (entry point with workspace %.12g)" % (e, er))


def test_noalloc_refuses_beyond_scratch():
    w = W.adjoint_morse(N=200000, L=60.0, seed=63)
    theta = np.array([[3.0, 6.0, 1.0], [1.0, 6.0, 1.0], [3.0, 6.0, 1.0]])
    rc, _, _ = noalloc_forces(w, [20, 20, 20], theta)
    assert rc == -6
    w = W.adjoint_morse(N=1000, L=160.0, seed=64)
    rc, _, _ = noalloc_forces(w, [60, 60, 60], theta)
    assert rc == -6
