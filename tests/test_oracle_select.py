"""Closed forms pinning the oracle's density selectors for alpha (CPU).

Contrast anchors at the decades where the piecewise-linear rule has known
values, the grayness of binary and uniform-gray densities, the Emin decade
table, and the compact-law features of a checkerboard (every neighbour
difference is the full span), a laminate (differences on two planes of one
axis) and a uniform field (no span).
"""
import math

import numpy as np
import pytest

from oracle import select as sel


def field_with_range(lo, hi, n=6, seed=0):
    a = np.random.default_rng(seed).uniform(lo, hi, (n, n, n))
    a.flat[0], a.flat[1] = lo, hi
    return a


@pytest.mark.parametrize('ratio,alpha', [(1.0, 0.0), (1e4, 0.5625), (1e5, 0.625), (1e7, 0.75), (1e12, 0.8125)])
def test_contrast_anchors(ratio, alpha):
    a = np.full((4, 4, 4), 0.5) if ratio == 1.0 else field_with_range(1.0 / ratio, 1.0)
    assert sel.contrast_alpha(a) == alpha


def test_grayness_and_gray_alpha():
    rng = np.random.default_rng(1)
    binary = (rng.random((6, 6, 6)) < .4).astype(float)
    assert sel.grayness(binary) == 0.0
    half = np.full((6, 6, 6), .5)
    assert abs(sel.grayness(half) - 1.0) < 1e-11
    a = field_with_range(1e-4, 1.0)
    # s = 4: base = b4; binary: 0.5625 + 0.075 g_ref -> 9/16; uniform gray: 0.5625 + 0.075 (g_ref - 1) -> 8/16
    assert sel.gray_alpha(a, binary) == 0.5625
    assert sel.gray_alpha(a, half) == 0.5


@pytest.mark.parametrize('emin,alpha', [(1e-4, 0.5625), (1e-6, 0.6875), (1e-9, 0.8125), (1e-5, 0.5625),
                                        (1e-8, 0.8125)])
def test_eta_table(emin, alpha):
    assert sel.eta_alpha(emin) == alpha


@pytest.mark.parametrize('lo', [1e-9, 1e-3])
def test_law_features_checkerboard(lo):
    n, hi = 8, 1.0
    z, y, x = np.indices((n, n, n))
    a = np.where((x + y + z) % 2 == 0, lo, hi)
    area, jump, contrast, complexity = sel.law_features(a)
    span = np.float32(np.log(np.float32(hi))) - np.float32(np.log(np.float32(lo)))
    assert area == 1.0 and jump == 1.0 and complexity == n
    # float32 as the reference: span / log(1e4) on a float32 tensor
    assert contrast == float(np.float32(span / np.float32(math.log(1e4))))
    expect = 1 - .75 / math.sqrt(contrast) - (2.5 / n) ** 8
    expect = 0.0 if expect < .375 else min(expect, 1.0)     # below 0.375 the law falls back to Green
    assert (expect > 0) == (lo == 1e-9)
    assert sel.compact8_alpha(a) == expect
    assert sel.compact8_alpha(a, quantize=True) == math.floor(16 * expect + .5) / 16


def test_law_features_laminate_and_uniform():
    n = 8
    a = np.ones((n, n, n))
    a[: n // 2] = 1e-2                                  # two interfaces normal to the first axis
    area, jump, contrast, complexity = sel.law_features(a)
    assert area == float(np.float32(2 / n / 3)) and jump == 1.0
    assert complexity == pytest.approx(2 / 3, rel=1e-7)
    # complexity 2/3 < 2.5: the complexity penalty drives the law to 0
    assert sel.compact8_alpha(a) == 0.0
    assert tuple(sel.law_features(np.full((4, 4, 4), .3))[[0, 2]]) == (0.0, 0.0)
    assert sel.compact8_alpha(np.full((4, 4, 4), .3)) == 0.0
