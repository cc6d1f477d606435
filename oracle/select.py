"""CPU restatement of the density selectors for the Green-Jacobi exponent alpha.

Test infrastructure only (the checker of ``pgj.select``). Formulas:
``contrast_alpha`` ``exp/asg_density_selector_20261105/run.py:28-33``,
``gray_alpha`` ``run.py:37-43``, ``eta_alpha`` ``run.py:52``,
``law_features`` ``solver/experiments/global_alpha_compact_law/law.py:9-18``
(FP32, as the reference evaluates it), ``compact8_alpha``
``exp/asg_gpu_resident_20261103/resident.py:163-168`` and ``run.py:60-66``.
Anchors: ``exp/asg_density_selector_20261105/MODELS_FROZEN.json``.
"""
import math

import numpy as np

ETA_ANCHORS = (0.5625, 0.6875, 0.8125)
CONTRAST_ANCHORS = (0.5625, 0.6875, 0.8125)
GRAY_ANCHORS = (0.5625, 0.75, 0.8125)
GRAY_GAMMA, GRAY_G_REF = 0.075, 0.3739899920170806


def _base(s, anchors):
    b4, b6, b9 = anchors
    if s < 4:
        return b4 * s / 4
    if s < 6:
        return b4 + (b6 - b4) * (s - 4) / 2
    return b6 + (b9 - b6) * (s - 6) / 3


def log10_contrast(a):
    a = np.asarray(a, float)
    return float(np.clip(np.log10(a.max() / a.min()), 0.0, 9.0))


def contrast_alpha(a, anchors=CONTRAST_ANCHORS):
    s = log10_contrast(a)
    return float(np.floor(np.clip(_base(s, anchors), 0, 1) * 16 + .5) / 16) if s > 1e-12 else 0.0


def grayness(rho):
    rho = np.asarray(rho, float)
    v = rho.mean()
    return float(np.clip((v - np.square(rho).mean()) / (v * (1 - v) + 1e-12), 0, 1))


def gray_alpha(a, rho, anchors=GRAY_ANCHORS, gamma=GRAY_GAMMA, g_ref=GRAY_G_REF):
    s = log10_contrast(a)
    value = _base(s, anchors) + gamma * (g_ref - grayness(rho))
    return float(np.floor(np.clip(value, 0, 1) * 16 + .5) / 16) if s > 1e-12 else 0.0


def eta_alpha(emin, anchors=ETA_ANCHORS):
    return float(anchors[int(np.argmin(np.abs(np.array([4, 6, 9]) + math.log10(emin))))])


def law_features(a):
    """(area, jump, contrast, complexity) of the field, computed and reduced in FP32 as the reference
    (``solver/experiments/global_alpha_compact_law/law.py:10-19``, float32 tensors and float32 means), then
    widened to FP64 (``resident.py:164``); axis 0 is the field's first axis."""
    f32 = np.float32
    ell = np.log(np.asarray(a, f32))
    lo, hi = ell.min(), ell.max()
    span = f32(hi - lo)
    q = (ell - lo) / np.maximum(span, f32(1e-12))
    n = q.shape[0]
    s1, s2, active = [], [], []
    for axis in range(3):
        d = np.abs(np.roll(q, -1, axis) - q)
        s1.append(d.mean(dtype=f32))
        s2.append(np.square(d).mean(dtype=f32))
        active.append((d > f32(1e-6)).astype(f32).mean(dtype=f32))
    total1 = f32(f32(s1[0] + s1[1]) + s1[2])
    total2 = f32(f32(s2[0] + s2[1]) + s2[2])
    area = f32(f32(f32(active[0] + active[1]) + active[2]) / f32(3))
    jump = f32(total2 / max(total1, f32(1e-20)))
    out = np.array([area, jump, f32(span / f32(math.log(10000.))), f32(area * f32(n))], dtype=f32)
    return out.astype(np.float64)


def compact8_alpha(a, quantize=False):
    area, jump, contrast, complexity = law_features(a)
    value = 1. - .75 * math.sqrt(max(area * jump, 0.)) / math.sqrt(max(contrast, 1e-30)) - \
        (2.5 / max(complexity, 1e-30)) ** 8
    value = min(max(value, 0.), 1.) if complexity > 1e-12 and contrast > 1e-12 else 0.
    value = 0. if value < .375 else value
    return math.floor(16 * value + .5) / 16 if quantize else value
