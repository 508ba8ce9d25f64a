"""FP64 reference of the differentiable overdamped trajectory.

Test infrastructure only: the product never imports this module.  It
restates, in torch float64 on the CPU, the map the GPU adjoint
differentiates, and takes exact gradients of it with torch autograd:

* pair set: all pairs with r^2 <= rc^2 under the minimum image (the cell
  loops of runner_dopair/runner_doself, src/mdcore/src/runner_dopair.cpp:
  91-340, 446-682), and s <= smax (potential_eval_ex's range test,
  src/mdcore/src/potential_eval.hpp:106-161);
* pair force: phi(r) d with phi = -E'(s) / max(r, FLT_EPSILON), per type
  pair either Morse, s = max(r, smin) - (R_i + R_j) + r0 + 1 (the shift rule
  of potential_create_morse, MxPotential.cpp:3395-3424), E(s) = D (u^2 - 2u),
  u = exp(-A (s - 1)), or Lennard-Jones 12-6, s = max(r, smin),
  E(s) = A s^-12 - B s^-6 (potential_create_LJ126, MxPotential.cpp:976-991)
  -- the closed forms of the reference's tables;
* bond force (bond_eval, src/mdcore/src/bond.cpp:99-295, on a
  potential_create_harmonic table): rb = clamp(r, bmin, bmax),
  phi = -E'(rb) / rb, E = K (rb - r0)^2;
* overdamped update (src/mdcore/src/engine_advance.cpp:126-239): per
  component dx = (dt F) / m clamped to +-max_dist_fraction * h, x += dx.

Its autograd gradients are pinned against FP64 central differences and the
force against -dE/dx in tests/test_adjoint_ref.py.
"""
import math

import numpy as np
import torch

FLT_EPSILON = float(np.finfo(np.float32).eps)


class RefSystem:
    def __init__(self, types, dim, cutoff, dt, mass=(1.0,), radius=(1.0,), cells=None, max_dist_fraction=0.05,
                 smin=1e-4, smax=None, kinds=None, bonds=None):
        self.types = torch.as_tensor(np.asarray(types), dtype=torch.long)
        self.N = len(self.types)
        self.ntypes = len(mass)
        self.dim = torch.as_tensor(np.asarray(dim, np.float64).reshape(3))
        if cells is None:
            cells = [max(3, int(math.floor(float(d) / cutoff))) for d in self.dim]
        self.h = self.dim / torch.as_tensor(np.asarray(cells, np.float64))
        self.maxx = (self.h * max_dist_fraction).to(torch.float32).to(torch.float64)
        self.cutoff = float(cutoff)
        self.dt = float(dt)
        self.imass = torch.as_tensor(1.0 / np.asarray(mass, np.float64))
        self.radius = torch.as_tensor(np.asarray(radius, np.float64))
        self.smin = float(smin)
        self.smax = float(cutoff if smax is None else smax)
        self.pair_of = torch.zeros((self.ntypes, self.ntypes), dtype=torch.long)
        k = 0
        for a in range(self.ntypes):
            for b in range(a, self.ntypes):
                self.pair_of[a, b] = self.pair_of[b, a] = k
                k += 1
        self.npairs = k
        # kind of each type pair: 0 Morse, 1 LJ 12-6
        self.kind = torch.zeros(k, dtype=torch.long)
        self.pmin = torch.full((k,), self.smin, dtype=torch.float64)
        self.pmax = torch.full((k,), self.smax, dtype=torch.float64)
        for (a, b), spec in (kinds or {}).items():
            # spec: kind, or (kind, smin, smax) for a domain of its own
            kd, lo, hi = (spec, self.smin, self.smax) if np.isscalar(spec) else spec
            self.kind[self.pair_of[a, b]] = int(kd)
            self.pmin[self.pair_of[a, b]] = float(lo)
            self.pmax[self.pair_of[a, b]] = float(hi)
        self.bonds = None
        if bonds is not None:
            self.bonds = (torch.as_tensor(np.asarray(bonds[0]), dtype=torch.long),
                          torch.as_tensor(np.asarray(bonds[1]), dtype=torch.long), float(bonds[2]), float(bonds[3]))

    def min_image(self, d):
        dim = self.dim.to(d.dtype)
        with torch.no_grad():
            k = torch.round(d / dim)
        return d - k * dim

    def pairs(self, x, rmax=None):
        """Index pairs i < j with r^2 <= rmax^2 (default: the cutoff)."""
        rmax = self.cutoff if rmax is None else rmax
        xs = x.detach()
        I, J = [], []
        for a in range(0, self.N, 512):
            d = self.min_image(xs[a:a + 512, None, :] - xs[None, :, :])
            r2 = (d * d).sum(-1)
            ii, jj = torch.nonzero(r2 <= rmax * rmax, as_tuple=True)
            ii = ii + a
            keep = ii < jj
            I.append(ii[keep])
            J.append(jj[keep])
        return torch.cat(I), torch.cat(J)

    def pair_terms(self, x, theta):
        """Pairs in range with E(s) and E'(s) of their potentials."""
        i, j = self.pairs(x)
        d = self.min_image(x[i] - x[j])
        r = torch.sqrt((d * d).sum(-1))
        p = self.pair_of[self.types[i], self.types[j]]
        P0, P1, r0 = theta[p, 0], theta[p, 1], theta[p, 2]
        lj = self.kind[p] == 1
        R = (self.radius[self.types[i]] + self.radius[self.types[j]]).to(x.dtype)
        pmin = self.pmin[p].to(x.dtype)
        re = torch.maximum(r, pmin)
        s = torch.where(lj, re, re - R + r0 + 1.0)
        with torch.no_grad():
            ok = s <= self.pmax[p].to(x.dtype)
        i, j, d, r, s, P0, P1, lj = i[ok], j[ok], d[ok], r[ok], s[ok], P0[ok], P1[ok], lj[ok]
        # Morse: D = P0, A = P1; LJ: A = P0, B = P1.  Each form is evaluated
        # on a harmless argument where the other applies, so that neither
        # puts a non-finite value into the gradient.
        sm = torch.where(lj, torch.ones_like(s), s)
        sl = torch.where(lj, s, torch.ones_like(s))
        u = torch.exp(-P1 * (sm - 1.0))
        is6 = sl ** -6
        e = torch.where(lj, P0 * is6 * is6 - P1 * is6, P0 * (u * u - 2.0 * u))
        ep = torch.where(lj, (-12.0 * P0 * is6 * is6 + 6.0 * P1 * is6) / sl, 2.0 * P1 * P0 * (u - u * u))
        return i, j, d, r, e, ep

    def bond_terms(self, x, theta):
        """Bonds with E(rb), E'(rb) and rb = clamp(r, bmin, bmax)."""
        i, j, bmin, bmax = self.bonds
        d = self.min_image(x[i] - x[j])
        r = torch.sqrt((d * d).sum(-1))
        rb = torch.clamp(r, min=bmin, max=bmax)
        K, r0 = theta[self.npairs, 0], theta[self.npairs, 1]
        return i, j, d, rb, K * (rb - r0) ** 2, 2.0 * K * (rb - r0)

    def energy(self, x, theta):
        e = self.pair_terms(x, theta)[4].sum()
        if self.bonds is not None:
            e = e + self.bond_terms(x, theta)[4].sum()
        return e

    def forces(self, x, theta):
        i, j, d, r, _, ep = self.pair_terms(x, theta)
        phi = -ep / torch.clamp(r, min=FLT_EPSILON)  # also runs in float32 to size FP32 rounding
        fp = phi[:, None] * d
        F = torch.zeros_like(x)
        F = F.index_add(0, i, fp).index_add(0, j, -fp)
        if self.bonds is not None:
            bi, bj, bd, rb, _, bep = self.bond_terms(x, theta)
            fb = (-bep / rb)[:, None] * bd
            F = F.index_add(0, bi, fb).index_add(0, bj, -fb)
        return F

    def step(self, x, theta):
        F = self.forces(x, theta)
        maxx = self.maxx.to(x.dtype)
        dx = (self.dt * F) * self.imass[self.types][:, None].to(x.dtype)
        dx = torch.maximum(torch.minimum(dx, maxx), -maxx)
        return x + dx

    def trajectory(self, x0, theta, nsteps):
        x = x0
        for _ in range(nsteps):
            x = self.step(x, theta)
        return x

    def histogram(self, x, nbins, rmax):
        """Soft pair-distance histogram, biweight bins (1 - t^2)^2."""
        i, j = self.pairs(x, rmax)
        d = self.min_image(x[i] - x[j])
        r = torch.sqrt((d * d).sum(-1))
        keep = (r < rmax).detach()
        i, j, r = i[keep], j[keep], r[keep]
        p = self.pair_of[self.types[i], self.types[j]]
        dr = rmax / nbins
        q = r / dr - 0.5
        k0 = torch.floor(q.detach()).long()
        npairs = self.ntypes * (self.ntypes + 1) // 2
        H = torch.zeros(npairs * nbins, dtype=x.dtype)
        for off in (0, 1):
            k = k0 + off
            t = q - k
            ok = ((k >= 0) & (k < nbins) & (t.abs() < 1)).detach()
            w = (1.0 - t * t) ** 2
            H = H.index_add(0, (p * nbins + k)[ok], w[ok])
        return H.reshape(npairs, nbins)

    def pair_counts(self):
        n = np.bincount(self.types.numpy(), minlength=self.ntypes).astype(np.float64)
        out = []
        for a in range(self.ntypes):
            for b in range(a, self.ntypes):
                out.append(n[a] * (n[a] - 1) / 2 if a == b else n[a] * n[b])
        return np.array(out)

    def rdf(self, x, nbins, rmax):
        H = self.histogram(x, nbins, rmax)
        dr = rmax / nbins
        r = (torch.arange(nbins, dtype=torch.float64) + 0.5) * dr
        vol = float(torch.prod(self.dim))
        pairs = torch.as_tensor(self.pair_counts())
        return H / (pairs[:, None] * 4.0 * math.pi * r[None, :] ** 2 * dr / vol)

    def loss(self, x0, theta, target, nsteps, nbins=64, rmax=2.5):
        xT = self.trajectory(x0, theta, nsteps)
        return ((self.rdf(xT, nbins, rmax) - target) ** 2).mean()
