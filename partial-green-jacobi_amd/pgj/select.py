"""On-device choice of the Green-Jacobi exponent alpha (the reference's density selectors).

``select_alpha`` writes alpha for a design into a device scalar with
``pgj_alpha_select``, so the scale can be built from it without a host round
trip (``pgj_resident_scale_dev``). Methods, as the reference names them:

* a number, ``'green'`` (0), ``'gj'`` (1) (``exp/asg_gpu_resident_20261103/resident.py:156-170``);
* ``'eta'``: the anchor of the Emin decade nearest to 1e-4, 1e-6, 1e-9
  (``exp/asg_density_selector_20261105/run.py:52``);
* ``'contrast'``: piecewise linear in log10(max a / min a), rounded to 1/16 (``run.py:28-33``);
* ``'gray'``: contrast plus a grayness term of the physical density (``run.py:37-43``);
* ``'compact8'`` / ``'compact8q'``: the compact law on FP32 features of a
  (``solver/experiments/global_alpha_compact_law/law.py:9-18``, ``resident.py:163-168``, ``run.py:60-66``).

The anchors are the frozen models of ``exp/asg_density_selector_20261105/MODELS_FROZEN.json``.
"""
import math

import numpy as np
import torch

from . import _lib

# MODELS_FROZEN.json "models": eta / contrast anchors (b4, b6, b9); gray anchors, gamma and g_ref
ETA_ANCHORS = (0.5625, 0.6875, 0.8125)
CONTRAST_ANCHORS = (0.5625, 0.6875, 0.8125)
GRAY_ANCHORS = (0.5625, 0.75, 0.8125)
GRAY_GAMMA = 0.075
GRAY_G_REF = 0.3739899920170806

_METHOD = {'green': 1, 'gj': 2, 'contrast': 4, 'gray': 5, 'compact8': 6, 'compact8q': 7}
FEATURES = ('log10_contrast', 'mean_rho', 'mean_rho2', 'area', 'jump', 'law_contrast', 'complexity')


def eta_alpha(emin, anchors=ETA_ANCHORS):
    """Anchor of the Emin decade (``run.py:52``): argmin |[4, 6, 9] + log10(Emin)|."""
    return float(anchors[int(np.argmin(np.abs(np.array([4, 6, 9]) + math.log10(emin))))])


class Selector:
    """Device workspace, feature and alpha buffers of one grid."""

    def __init__(self, grid, device='cuda'):
        self.grid = grid
        self.ws = torch.empty(_lib.lib().pgj_alpha_ws_bytes(grid) // 8 + 1, dtype=torch.float64, device=device)
        self.features = torch.zeros(len(FEATURES), dtype=torch.float64, device=device)
        self.alpha = torch.zeros((), dtype=torch.float64, device=device)

    def __call__(self, a, method, emin=None, rho=None):
        """alpha of the SIMP field ``a`` (device) into ``self.alpha``; no host synchronisation.

        ``emin`` is needed by ``'eta'``, the physical density ``rho`` by ``'gray'``. Returns ``self.alpha``.
        """
        if isinstance(method, str) and method == 'eta':
            if emin is None:
                raise ValueError("the 'eta' selector needs emin")
            code, params = 0, [eta_alpha(emin)]
        elif isinstance(method, str):
            if method not in _METHOD:
                raise ValueError(f'unknown alpha selector {method!r}')
            code = _METHOD[method]
            params = {4: list(CONTRAST_ANCHORS), 5: list(GRAY_ANCHORS) + [GRAY_GAMMA, GRAY_G_REF]}.get(code, [0.0])
            if code == 5 and rho is None:
                raise ValueError("the 'gray' selector needs the physical density rho")
        else:
            code, params = 0, [float(method)]
            if not 0.0 <= params[0] <= 1.0:
                raise ValueError(f'alpha must be in [0, 1], got {params[0]}')
        a = a.reshape(-1).contiguous()
        r = None if rho is None else rho.reshape(-1).contiguous()
        _lib.call('pgj_alpha_select', self.grid, code, _lib.host_doubles(params), _lib.ptr(a),
                  None if r is None else _lib.ptr(r), _lib.ptr(self.ws), _lib.ptr(self.features),
                  _lib.ptr(self.alpha), _lib.stream())
        return self.alpha
