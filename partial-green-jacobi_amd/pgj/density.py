"""Seeded synthetic densities for the workloads (input generation, host numpy).

rho = lo + (1 - lo) (f - min f) / (max f - min f), where f is the periodic
(2r+1)^3 box mean of U(0,1) noise from numpy's default_rng(seed), axes (z, y, x).
The reference has no such recipe; this defines it for the benchmark workloads.
"""
import numpy as np


def filtered_noise(n, seed, radius, lo=1e-3):
    noise = np.random.default_rng(seed).random((n, n, n))
    # separable periodic box sum (exact in the order z, y, x)
    f = noise
    for axis in range(3):
        acc = np.zeros_like(f)
        for d in range(-radius, radius + 1):
            acc += np.roll(f, d, axis=axis)
        f = acc
    f = f / (2 * radius + 1) ** 3
    return lo + (1 - lo) * (f - f.min()) / (f.max() - f.min())


def simp(rho, emin, p=3.0):
    """a = Emin + rho^p (1 - Emin)."""
    return emin + (1 - emin) * np.asarray(rho, dtype=np.float64) ** p
