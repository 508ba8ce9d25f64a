"""Differentiable overdamped trajectories on the GPU.

``MorseSystem`` wraps the C ABI's adjoint object (include/mxg.h,
``mxg_adjoint_*``); ``trajectory`` and ``soft_rdf`` are torch.autograd
Functions whose backward passes are the hand-written HIP adjoint kernels:

* ``trajectory(system, x0, theta, nsteps, every)`` runs ``nsteps``
  overdamped steps (engine_step's overdamped update,
  src/mdcore/src/engine_advance.cpp:126-239, with the pair force of
  potential_eval_ex, src/mdcore/src/potential_eval.hpp:106-161, and the bond
  force of bond_eval, src/mdcore/src/bond.cpp:99-295, in closed form) and
  keeps the state every ``every`` steps; its backward recomputes each
  segment from its checkpoint and sweeps it in reverse, returning dL/dx0 and
  dL/dtheta accumulated in FP64.
* ``soft_rdf(system, x, nbins, rmax)`` is the per-type-pair pair-distance
  histogram with biweight bins, and ``rdf_normalise`` turns it into g(r).

theta holds one row per type pair (ti <= tj, row-major upper triangle): the
arguments of potential_create_morse (d, a, r0) for a Morse pair or of
potential_create_LJ126 (A, B, unused) for a Lennard-Jones pair; with bonds,
one more row holds the harmonic bond's (K, r0, unused)
(potential_create_harmonic).  torch is used for device memory and the
current stream only.
"""
import ctypes as C
import math

import numpy as np
import torch

from . import _capi


class AdjointError(RuntimeError):
    pass


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream(device):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


MORSE, LJ = 0, 1


class MorseSystem:
    """N particles of up to four types in a periodic box, overdamped, with
    Morse or LJ 12-6 pairs (kinds: {(ti, tj): MORSE or LJ, or (kind, smin,
    smax) with a domain of its own}; default Morse on [smin, smax]) and
    optional harmonic bonds (bonds: (i, j, rmin, rmax))."""

    def __init__(self, types, dim, cutoff, dt, mass=(1.0,), radius=(1.0,), cells=None, max_dist_fraction=0.05,
                 smin=1e-4, smax=None, device="cuda", kinds=None, bonds=None):
        L = _capi.lib()
        types = np.ascontiguousarray(types, dtype=np.int32)
        self.N = int(types.shape[0])
        self.ntypes = len(mass)
        self.npairs = self.ntypes * (self.ntypes + 1) // 2
        self.dim = np.ascontiguousarray(dim, dtype=np.float64).reshape(3)
        if cells is None:
            cells = [max(3, int(math.floor(d / cutoff))) for d in self.dim]
        self.cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(3)
        self.cutoff = float(cutoff)
        self.dt = float(dt)
        self.smin = float(smin)
        self.smax = float(cutoff if smax is None else smax)
        self.types = types
        self.device = torch.device(device)
        m = np.ascontiguousarray(mass, dtype=np.float64)
        r = np.ascontiguousarray(radius, dtype=np.float64)
        h = C.c_void_p()
        rc = L.mxg_adjoint_create(C.byref(h), self.N, self.dim.ctypes.data_as(_capi.c_double_p),
                                  self.cells.ctypes.data_as(_capi.c_int_p), self.cutoff, self.dt,
                                  float(max_dist_fraction), self.ntypes, m.ctypes.data_as(_capi.c_double_p),
                                  r.ctypes.data_as(_capi.c_double_p), types.ctypes.data_as(_capi.c_int_p))
        self._h = h
        self._check(rc)
        self.nbonds = 0
        for (a, b), spec in (kinds or {}).items():
            # spec: kind, or (kind, smin, smax) for a domain of its own
            k, lo, hi = (spec, 0.0, 0.0) if np.isscalar(spec) else spec
            self._check(L.mxg_adjoint_set_pair_kind(self._h, int(a), int(b), int(k), float(lo), float(hi)))
        if bonds is not None:
            bi = np.ascontiguousarray(bonds[0], np.int32)
            bj = np.ascontiguousarray(bonds[1], np.int32)
            self._check(L.mxg_adjoint_set_bonds(self._h, len(bi), bi.ctypes.data_as(_capi.c_int_p),
                                                bj.ctypes.data_as(_capi.c_int_p), float(bonds[2]),
                                                float(bonds[3])))
            self.nbonds = len(bi)
        self.nrows = self.npairs + (1 if self.nbonds else 0)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.lib().mxg_adjoint_free(h)
            except Exception:
                pass
            self._h = None

    def _check(self, rc):
        if rc is not None and rc < 0:
            msg = _capi.lib().mxg_adjoint_error_message(self._h)
            raise AdjointError("%d: %s" % (rc, msg.decode() if msg else ""))

    def _theta(self, theta):
        th = np.ascontiguousarray(torch.as_tensor(theta).detach().to("cpu", torch.float64).numpy()).reshape(-1)
        if th.size != 3 * self.nrows:
            raise ValueError("theta needs %d rows of 3 (type pairs, then the bond row)" % self.nrows)
        return th

    def _x(self, x):
        x = torch.as_tensor(x)
        if x.shape != (self.N, 3):
            raise ValueError("positions must have shape (%d, 3)" % self.N)
        return x.detach().to(self.device, torch.float64).contiguous()

    # raw calls ------------------------------------------------------------
    def run(self, x0, theta, nsteps, every=10):
        """Forward only: returns (xT, checkpoints)."""
        th = self._theta(theta)
        x0 = self._x(x0)
        nck = (nsteps + every - 1) // every
        ckpt = torch.empty((nck, self.N, 4), dtype=torch.float32, device=self.device)
        xT = torch.empty((self.N, 3), dtype=torch.float64, device=self.device)
        self._check(_capi.lib().mxg_adjoint_forward(
            self._h, th.ctypes.data_as(_capi.c_double_p), self.smin, self.smax, _ptr(x0), _ptr(xT), _ptr(ckpt),
            int(nsteps), int(every), _stream(self.device)))
        return xT, ckpt

    def adjoint(self, theta, ckpt, nsteps, every, gxT):
        """Backward sweep: returns (dL/dx0, dL/dtheta) in FP64."""
        th = self._theta(theta)
        gxT = self._x(gxT)
        gx0 = torch.empty((self.N, 3), dtype=torch.float64, device=self.device)
        gth = np.zeros(3 * self.nrows, np.float64)
        self._check(_capi.lib().mxg_adjoint_backward(
            self._h, th.ctypes.data_as(_capi.c_double_p), self.smin, self.smax, _ptr(ckpt), int(nsteps), int(every),
            _ptr(gxT), _ptr(gx0), gth.ctypes.data_as(_capi.c_double_p), _stream(self.device)))
        return gx0, torch.from_numpy(gth.reshape(self.nrows, 3))

    def forces(self, x, theta):
        """Forces (N, 3) FP32 and the total pair and bond energy (FP64 sum)."""
        th = self._theta(theta)
        x = self._x(x)
        f = torch.empty((self.N, 3), dtype=torch.float32, device=self.device)
        e = C.c_double()
        self._check(_capi.lib().mxg_adjoint_forces(
            self._h, th.ctypes.data_as(_capi.c_double_p), self.smin, self.smax, _ptr(x), _ptr(f), C.byref(e),
            _stream(self.device)))
        return f, e.value

    def histogram(self, x, nbins, rmax):
        x = self._x(x)
        hist = torch.empty((self.npairs, nbins), dtype=torch.float64, device=self.device)
        self._check(_capi.lib().mxg_adjoint_rdf(self._h, _ptr(x), int(nbins), float(rmax), _ptr(hist),
                                                _stream(self.device)))
        return hist

    def histogram_backward(self, x, nbins, rmax, ghist):
        x = self._x(x)
        ghist = ghist.detach().to(self.device, torch.float64).contiguous()
        gx = torch.empty((self.N, 3), dtype=torch.float64, device=self.device)
        self._check(_capi.lib().mxg_adjoint_rdf_backward(self._h, _ptr(x), int(nbins), float(rmax), _ptr(ghist),
                                                         _ptr(gx), _stream(self.device)))
        return gx

    def pair_counts(self):
        """Number of particle pairs per type pair (ti <= tj)."""
        n = np.bincount(self.types, minlength=self.ntypes).astype(np.float64)
        out = []
        for a in range(self.ntypes):
            for b in range(a, self.ntypes):
                out.append(n[a] * (n[a] - 1) / 2 if a == b else n[a] * n[b])
        return np.array(out)


class _Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, theta, system, nsteps, every):
        xT, ckpt = system.run(x0, theta, nsteps, every)
        ctx.system, ctx.nsteps, ctx.every = system, nsteps, every
        ctx.theta = theta.detach().clone()
        ctx.x0_meta = (x0.dtype, x0.device)
        ctx.save_for_backward(ckpt)
        return xT

    @staticmethod
    def backward(ctx, gxT):
        (ckpt,) = ctx.saved_tensors
        gx0, gth = ctx.system.adjoint(ctx.theta, ckpt, ctx.nsteps, ctx.every, gxT)
        dt, dev = ctx.x0_meta
        return gx0.to(dev, dt), gth.to(ctx.theta.device, ctx.theta.dtype), None, None, None


class _SoftRDF(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, system, nbins, rmax):
        ctx.system, ctx.nbins, ctx.rmax = system, nbins, rmax
        ctx.x_meta = (x.dtype, x.device)
        ctx.save_for_backward(x.detach())
        return system.histogram(x, nbins, rmax)

    @staticmethod
    def backward(ctx, ghist):
        (x,) = ctx.saved_tensors
        gx = ctx.system.histogram_backward(x, ctx.nbins, ctx.rmax, ghist)
        dt, dev = ctx.x_meta
        return gx.to(dev, dt), None, None, None


def trajectory(system, x0, theta, nsteps, every=10):
    """Final positions (N, 3) FP64 after nsteps steps; differentiable in x0
    and theta (nrows, 3): the type pairs, then the bond row."""
    return _Trajectory.apply(x0, theta, system, int(nsteps), int(every))


def soft_rdf(system, x, nbins=64, rmax=2.5):
    """Soft pair-distance histogram (npairs, nbins); differentiable in x."""
    return _SoftRDF.apply(x, system, int(nbins), float(rmax))


def rdf_normalise(hist, system, rmax):
    """g(r) per type pair: histogram / (pairs * shell volume / box volume)."""
    nbins = hist.shape[-1]
    dr = rmax / nbins
    r = (torch.arange(nbins, dtype=torch.float64, device=hist.device) + 0.5) * dr
    vol = float(np.prod(system.dim))
    pairs = torch.as_tensor(system.pair_counts(), dtype=torch.float64, device=hist.device)
    ideal = pairs[:, None] * (4.0 * math.pi * r[None, :] ** 2 * dr) / vol
    return hist / ideal


def rdf_loss(system, x0, theta, target, nsteps, every=10, nbins=64, rmax=2.5):
    """Mean-squared deviation of the final g(r) from target (npairs, nbins)."""
    xT = trajectory(system, x0, theta, nsteps, every)
    g = rdf_normalise(soft_rdf(system, xT, nbins, rmax), system, rmax)
    return ((g - target) ** 2).mean()
