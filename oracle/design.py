"""Densities, C^H, sensitivities and OC updates in numpy (test infrastructure only).

- ``filtered_noise``: the seeded density recipe of the workloads (uniform
  noise, periodic box mean of radius r, affine rescale to [1e-3, 1]); it has
  no reference counterpart and is pinned by its definition.
- ``effective_tensor``: ``exp/asg_sctop_reproduction_20261105/reproduction.py:166-176``.
- ``shear_objective``: ``reproduction.py:153-164`` ('shear').
- ``PaperOC``: ``exp/asg_sctop_reproduction_20261105/outer.cu:8-39``,
  ``reproduction.py:69-118``, CPU form ``verify.py:12-37``, with volume,
  move and Emin as parameters (the reference fixes move = 0.1).
- ``GaussianOC``: ``exp/asg_gpu_resident_20261103/outer.{py,cu}``, CPU form
  ``exp/asg_topopt_20261103/study.py:61-100``, with volume, move and Emin as
  parameters (the reference fixes 0.3, 0.1, 1e-4).
Pinned in ``tests/test_oracle_design.py`` (solid C^H = C0, laminate closed
form, finite-difference sensitivities, dense filter matrix, scipy
``gaussian_filter(mode='wrap')``).
"""
import math

import numpy as np
from scipy.ndimage import gaussian_filter

from . import hex8


def filtered_noise(n, seed, radius, lo=1e-3):
    """rho = lo + (1-lo)(f - min)/(max - min), f = periodic (2r+1)^3 box mean of U(0,1) noise."""
    noise = np.random.default_rng(seed).random((n, n, n))
    total = np.zeros_like(noise)
    for dz in range(-radius, radius + 1):
        for dy in range(-radius, radius + 1):
            for dx in range(-radius, radius + 1):
                total += np.roll(noise, shift=(dz, dy, dx), axis=(0, 1, 2))
    f = total / (2 * radius + 1) ** 3
    return lo + (1 - lo) * (f - f.min()) / (f.max() - f.min())


def simp(rho, emin, p=3.0):
    """a = Emin + rho^p (1 - Emin) (``outer.cu:20-28``)."""
    return emin + (1 - emin) * np.asarray(rho, float) ** p


def effective_tensor(a, solutions, C, voxel=None):
    """C^H by energy polarization (``reproduction.py:166-176``)."""
    a = np.asarray(a, float)
    loads = np.eye(6)
    e = hex8.energies(solutions, a.shape, loads, C, voxel)
    d = (e * a.reshape(1, -1)).mean(1)
    H = np.diag(d)
    for i in range(6):
        for j in range(i + 1, 6):
            pair = hex8.energies(solutions[i:i + 1] + solutions[j:j + 1], a.shape,
                                 (loads[i] + loads[j])[None], C, voxel)
            q = (pair * a.reshape(1, -1)).mean()
            H[i, j] = H[j, i] = 0.5 * (q - d[i] - d[j])
    return H


def shear_objective(a, solutions, C, voxel=None):
    """(objective, local, diagonal) for the 'shear' objective (``reproduction.py:153-164``)."""
    a = np.asarray(a, float)
    e = hex8.energies(solutions, a.shape, np.eye(6), C, voxel)
    diagonal = (e * a.reshape(1, -1)).mean(1)
    local = -e[3:].sum(0)
    return float((local * a.reshape(-1)).mean()), local.reshape(a.shape), diagonal


def cone_stencil(rmin):
    """Offsets (m,3) and weights of the clipped cone filter, in reference loop order (``reproduction.py:38-45``)."""
    radius = math.ceil(rmin) - 1
    offsets, weights = [], []
    for dz in range(-radius, radius + 1):
        for dx in range(-radius, radius + 1):
            for dy in range(-radius, radius + 1):
                w = rmin - math.sqrt(dx * dx + dy * dy + dz * dz)
                if w > 0:
                    offsets.append([dx, dy, dz])
                    weights.append(w)
    return np.array(offsets, dtype=np.int32), np.array(weights)


def cone_filter(x, offsets, weights):
    """Non-periodic clipped cone filter sum_j w_j x[i + o_j], in stencil order (``outer.cu:8-19``)."""
    x = np.asarray(x, float)
    n = x.shape[0]
    out = np.zeros_like(x)
    zz, yy, xx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing='ij')
    for (dx, dy, dz), w in zip(offsets, weights):
        xj, yj, zj = xx + dx, yy + dy, zz + dz
        ok = (xj >= 0) & (xj < n) & (yj >= 0) & (yj < n) & (zj >= 0) & (zj < n)
        vals = np.where(ok, x[np.clip(zj, 0, n - 1), np.clip(yj, 0, n - 1), np.clip(xj, 0, n - 1)], 0.0)
        out = out + np.where(ok, w * vals, 0.0)
    return out


class PaperOC:
    """Clipped cone filter, eta-projection and exponent-0.3 OC (``reproduction.py:69-118``)."""

    def __init__(self, n, volume, rmin=3.0, beta=4.0, eta=0.5, emin=1e-9, move=0.1, max_bisect=180):
        self.n, self.volume, self.beta, self.eta, self.emin = n, volume, beta, eta, emin
        self.move, self.max_bisect = move, max_bisect
        self.offsets, self.weights = cone_stencil(rmin)
        self.hs = self.filter(np.ones((n, n, n)))

    def filter(self, x):
        return cone_filter(x, self.offsets, self.weights)

    def project(self, t):
        b, eta = self.beta, self.eta
        den = np.tanh(b * eta) + np.tanh(b * (1 - eta))
        q = np.tanh(b * (t / self.hs - eta))
        rho = (np.tanh(b * eta) + q) / den
        return rho, b * (1 - q * q) / den, self.emin + (1 - self.emin) * rho * rho * rho

    def physical(self, x):
        return self.project(self.filter(x))

    def update(self, x, ph, dp, energy):
        """Returns (new, bisections, g, dv); ``energy`` is the local objective density."""
        n = self.n
        g = self.filter(dp * (3 * (1 - self.emin) * ph * ph) * energy / self.hs) / n ** 3
        dv = self.filter(dp / self.hs)
        ratio = -g / dv
        lo, hi, count = 0.0, 1e9, 0
        new = x.copy()
        while True:
            mid = 0.5 * (lo + hi)
            v = x * np.maximum(1e-10, ratio / mid) ** 0.3
            new = np.maximum(0.001, np.maximum(x - self.move, np.minimum(1.0, np.minimum(x + self.move, v))))
            rho, _, _ = self.physical(new)
            if rho.mean() > self.volume:
                lo = mid
            else:
                hi = mid
            count += 1
            if not ((hi - lo) / (hi + lo) > 1e-4 and hi > 1e-40) or count >= self.max_bisect:
                break
        return new, count, g, dv


def gaussian_weights(sigma, radius):
    """scipy's 1-D Gaussian kernel: exp(-0.5/sigma^2 x^2) normalized."""
    x = np.arange(-radius, radius + 1)
    phi = np.exp(-0.5 / (sigma * sigma) * x ** 2)
    return phi / phi.sum()


class GaussianOC:
    """Periodic Gaussian filter, tanh projection, 45-step sqrt OC (``outer.py:12-59``, ``study.py:61-100``)."""

    def __init__(self, n, volume=0.3, move=0.1, emin=1e-4, p=3.0, sigma_per_32=0.9, bisections=45):
        self.n, self.volume, self.move, self.emin, self.p = n, volume, move, emin, p
        self.sigma = sigma_per_32 * n / 32
        self.bisections = bisections

    def filter(self, x):
        return gaussian_filter(np.asarray(x, float), self.sigma, mode='wrap')

    def physical(self, x, beta):
        t = self.filter(x)
        den = 2 * np.tanh(beta / 2)
        q = np.tanh(beta * (t - 0.5))
        ph = (np.tanh(beta / 2) + q) / den
        dp = beta * (1 - q ** 2) / den
        return ph, dp, self.emin + (1 - self.emin) * ph ** self.p

    def update(self, x, ph, dp, energy, beta):
        g = self.filter(dp * self.p * (1 - self.emin) * ph ** 2 * energy)
        dv = self.filter(dp)
        lower, upper = np.maximum(0.0, x - self.move), np.minimum(1.0, x + self.move)
        ratio = np.maximum(g, 1e-30) / np.maximum(dv, 1e-30)
        lo, hi = 0.0, float(ratio.max()) * 1e4
        cand = x
        for _ in range(self.bisections):
            lam = (lo + hi) / 2
            cand = np.clip(x * np.sqrt(ratio / max(lam, 1e-30)), lower, upper)
            if self.physical(cand, beta)[0].mean() > self.volume:
                lo = lam
            else:
                hi = lam
        return cand, g, dv
