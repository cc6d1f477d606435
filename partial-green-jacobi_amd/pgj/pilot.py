"""Design-loop constants and initial designs, as the reference's callers use them.

``CONFIG`` holds the frozen SIMP/OC configuration of
``exp/asg_topopt_20261103/study.py:22-25`` that the design loop reads
(``examples/shear_design.py:43``: ``beta_stages[min(k // 25, 3)]``), and
``initial(n, seed)`` restates ``study.py:74-87``: twelve random periodic cosine
modes scaled to standard deviation 0.12, shifted by a 50-step bisection so that
the beta = 1 physical density has mean 0.3. The physical density there uses
scipy's periodic Gaussian filter (``study.py:55-62``); this is host-side input
generation, as in the reference. ``paper_initial(n, field)`` restates the
Study-II initial fields of ``exp/asg_sctop_reproduction_20261105/reproduction.py:21-34``
(spherical voids of density 0.001 in a 0.4 background at the MATLAB-style
centres).
"""
import math
import time

import numpy as np
import torch
from scipy.ndimage import gaussian_filter

CONFIG = dict(volume=.30, eta=1e-4, p=3., move=.10, filter_sigma_per_32=.9, beta_stages=[1., 2., 4., 8.],
              stage_steps=25, max_steps=180, rtol=1e-8)


def physical(x, beta):
    """(ph, dp) of the Gaussian-filtered tanh projection (``study.py:55-62``)."""
    t = gaussian_filter(x, CONFIG['filter_sigma_per_32'] * x.shape[0] / 32, mode='wrap')
    den = 2 * np.tanh(beta / 2)
    q = np.tanh(beta * (t - .5))
    return (np.tanh(beta / 2) + q) / den, beta * (1 - q ** 2) / den


def initial(n, seed):
    """Seeded Study-I initial design (n, n, n) in [0.01, 0.99] with beta = 1 physical volume 0.3."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*([(np.arange(n) + .5) / n] * 3), indexing='ij')
    field = np.zeros((n, n, n))
    for _ in range(12):
        k = rng.integers(-2, 3, 3)
        if not np.any(k):
            k[0] = 1
        field += rng.normal() * np.cos(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + rng.uniform(0, 2 * np.pi))
    field = .12 * field / field.std()
    lo, hi = -1., 1.
    for _ in range(50):
        shift = (lo + hi) / 2
        candidate = np.clip(.3 + field + shift, .01, .99)
        if physical(candidate, 1.)[0].mean() > .3:
            hi = shift
        else:
            lo = shift
    return candidate


def beta(step):
    """Projection sharpness of design step ``step`` (0-based)."""
    return CONFIG['beta_stages'][min(step // CONFIG['stage_steps'], len(CONFIG['beta_stages']) - 1)]


def sync():
    torch.cuda.synchronize()
    return time.perf_counter()


def paper_initial(n, field):
    """Study-II initial design ``field`` in 1..4 (axes z, y, x; 1-based centres as the reference)."""
    z, y, x = np.meshgrid(*([np.arange(1, n + 1, dtype=float)] * 3), indexing='ij')
    if field == 1:
        centers, radius = [(n + 1) / 2], n / 3
    elif field == 2:
        centers, radius = [1, (n + 1) / 2, n], math.floor(n / 5)
    elif field == 3:
        d = math.ceil((n - 1) / 6)
        centers, radius = [1, (n + 1) / 2 - d, (n + 1) / 2 + d, n], math.floor(n / 7)
    elif field == 4:
        centers, radius = np.linspace(1, n, 5), math.floor(n / 12 + .5)
    else:
        raise ValueError(field)
    dx = np.min((x[..., None] - centers) ** 2, axis=-1)
    dy = np.min((y[..., None] - centers) ** 2, axis=-1)
    dz = np.min((z[..., None] - centers) ** 2, axis=-1)
    return np.where(dx + dy + dz < radius ** 2, .001, .4)
