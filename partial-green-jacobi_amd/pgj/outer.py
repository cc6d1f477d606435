"""Device-resident filters, projections and OC density updates on libpgj.

``PaperOuter`` mirrors ``exp/asg_sctop_reproduction_20261105/reproduction.py:69-118``
(clipped cone filter, eta-projection, exponent-0.3 OC with relative-tolerance
bisection) and ``GaussianOuter`` mirrors ``exp/asg_gpu_resident_20261103/outer.py:12-59``
(periodic Gaussian, tanh projection, 45-step sqrt OC). Volume, move limit and
Emin are parameters (the reference fixes move 0.1, and Study I also volume 0.3
and Emin 1e-4). The bisection runs in captured graph chunks; a chunk after the
stop condition is a no-op, so chunking does not change results.
"""
import math

import numpy as np
import torch

from . import _lib


def cone_stencil(rmin):
    """Offsets (m, 3) [dx, dy, dz] and weights rmin - |d| > 0, reference loop order (``reproduction.py:38-45``)."""
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


def gaussian_weights(sigma, radius):
    """1-D Gaussian taps exp(-0.5/sigma^2 x^2), normalized (scipy.ndimage's kernel)."""
    x = np.arange(-radius, radius + 1)
    phi = np.exp(-0.5 / (sigma * sigma) * x ** 2)
    return phi / phi.sum()


class _Fields:
    def _alloc(self, n, names, device):
        for name in names:
            setattr(self, name, torch.empty((n, n, n), dtype=torch.float64, device=device))
        self.ws = torch.empty(_lib.lib().pgj_oc_ws_bytes(n) // 8 + 1, dtype=torch.float64, device=device)
        self.change = torch.zeros(1, dtype=torch.float64, device=device)

    def set_initial(self, x):
        self.x.copy_(torch.as_tensor(x, dtype=torch.float64).reshape(self.x.shape))

    def advance(self):
        self.x.copy_(self.new)

    def _change(self):
        _lib.call('pgj_absmax_diff', self.n, _lib.ptr(self.new), _lib.ptr(self.x), _lib.ptr(self.ws),
                  _lib.ptr(self.change), _lib.stream())


class PaperOuter(_Fields):
    """Cone filter + eta-projection + exponent-0.3 OC (``reproduction.py:69-118``)."""

    def __init__(self, n, volume, rmin=3.0, beta=4.0, eta=0.5, emin=1e-9, move=0.1, max_bisect=180, chunk=12,
                 device='cuda'):
        self.n, self.volume, self.beta, self.eta, self.emin, self.move = n, volume, beta, eta, emin, move
        self.max_bisect, self.chunk = max_bisect, chunk
        off, w = cone_stencil(rmin)
        self.offsets = torch.tensor(off, device=device).contiguous()
        self.weights = torch.tensor(w, dtype=torch.float64, device=device)
        self._alloc(n, ['x', 'ph', 'dp', 'a', 'energy', 'new', 'g', 'dv', 'ratio', 't', 'hs', 'trial_ph',
                        'trial_dp', 'trial_a', 'tmp'], device)
        self.bounds = torch.zeros(4, dtype=torch.float64, device=device)
        self.raw_filter(torch.ones_like(self.x), self.hs)
        self.physical_graph = self.prepare_graph = self.bisect_graph = None

    def raw_filter(self, x, out, conditional=False):
        _lib.call('pgj_cone_filter', self.n, _lib.ptr(x.contiguous()), _lib.ptr(self.offsets), _lib.ptr(self.weights),
                  len(self.weights), _lib.ptr(out), _lib.ptr(self.bounds), int(conditional), _lib.stream())

    def project(self, t, ph, dp, a, conditional=False, volume=False):
        _lib.call('pgj_cone_project', self.n, _lib.ptr(t), _lib.ptr(self.hs), _lib.ptr(ph), _lib.ptr(dp), _lib.ptr(a),
                  _lib.ptr(self.bounds), self.beta, self.eta, self.emin, int(conditional),
                  _lib.ptr(self.ws) if volume else None, _lib.stream())

    def physical(self):
        self.raw_filter(self.x, self.t)
        self.project(self.t, self.ph, self.dp, self.a)

    def _prepare(self):
        _lib.call('pgj_cone_sensitivity', self.n, _lib.ptr(self.dp), _lib.ptr(self.ph), _lib.ptr(self.energy),
                  _lib.ptr(self.hs), self.emin, _lib.ptr(self.tmp), _lib.ptr(self.new), _lib.stream())
        self.raw_filter(self.tmp, self.g)
        self.raw_filter(self.new, self.dv)
        _lib.call('pgj_cone_ratio', self.n, _lib.ptr(self.g), _lib.ptr(self.dv), _lib.ptr(self.ratio),
                  _lib.ptr(self.bounds), _lib.stream())

    def _bisect_chunk(self):
        for _ in range(self.chunk):
            _lib.call('pgj_cone_candidate', self.n, _lib.ptr(self.x), _lib.ptr(self.ratio), _lib.ptr(self.bounds),
                      self.move, _lib.ptr(self.new), _lib.stream())
            self.raw_filter(self.new, self.t, True)
            self.project(self.t, self.trial_ph, self.trial_dp, self.trial_a, True, True)
            _lib.call('pgj_cone_bisect', self.n, _lib.ptr(self.ws), _lib.ptr(self.bounds), self.volume,
                      _lib.stream())

    def update(self):
        """Sensitivity filtering and the bisection; returns the bisection count (``reproduction.py:91-104``)."""
        if self.bisect_graph is None:
            self._prepare()
            rounds = 0
            while rounds < self.max_bisect:
                self._bisect_chunk()
                rounds += self.chunk
                if float(self.bounds[2]) == 0.0:
                    break
        else:
            self.prepare_graph.replay()
            rounds = 0
            while rounds < self.max_bisect:
                self.bisect_graph.replay()
                rounds += self.chunk
                if float(self.bounds[2]) == 0.0:
                    break
        self._change()
        return int(self.bounds[3])

    def capture(self):
        self.x.fill_(.4)
        self.energy.fill_(-1.)
        self.physical()
        self._prepare()
        self._bisect_chunk()
        torch.cuda.synchronize()
        for name, fn in (('physical_graph', self.physical), ('prepare_graph', self._prepare),
                         ('bisect_graph', self._bisect_chunk)):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                fn()
            setattr(self, name, g)
        torch.cuda.synchronize()
        return self

    def evaluate(self):
        if self.physical_graph is None:
            self.physical()
        else:
            self.physical_graph.replay()

    def step(self, e):
        self.energy.copy_(torch.as_tensor(e).reshape(self.energy.shape))
        return self.update()


class GaussianOuter(_Fields):
    """Periodic Gaussian + tanh projection + 45-step sqrt OC (``exp/asg_gpu_resident_20261103/outer.py:12-59``)."""

    def __init__(self, n, volume=0.3, move=0.1, emin=1e-4, penal=3.0, sigma_per_32=0.9, bisections=45,
                 device='cuda'):
        self.n, self.volume, self.move, self.emin, self.penal = n, volume, move, emin, penal
        self.bisections = bisections
        sigma = sigma_per_32 * n / 32
        self.radius = int(4 * sigma + .5)
        if self.radius > 4 * n:
            raise ValueError('Gaussian radius exceeds the supported range')
        self.weights = torch.tensor(gaussian_weights(sigma, self.radius), dtype=torch.float64, device=device)
        self._alloc(n, ['x', 'ph', 'dp', 'a', 'energy', 'new', 'g', 'dv', 'ratio', 'tmp1', 'tmp2', 't', 'trial_ph',
                        'trial_dp', 'trial_a', 'gin'], device)
        self.beta = 1.0
        self.bounds = torch.zeros(2, dtype=torch.float64, device=device)
        self.physical_graph = self.update_graph = None

    def gaussian(self, x, out):
        _lib.call('pgj_gauss_filter', self.n, _lib.ptr(x.contiguous()), _lib.ptr(self.weights), self.radius,
                  _lib.ptr(self.tmp1), _lib.ptr(self.tmp2), _lib.ptr(out), _lib.stream())

    def _project(self, t, ph, dp, a, beta, volume=False):
        _lib.call('pgj_gauss_project', self.n, _lib.ptr(t), float(beta), _lib.ptr(ph), _lib.ptr(dp), _lib.ptr(a),
                  self.emin, _lib.ptr(self.ws) if volume else None, _lib.stream())

    def physical(self, beta):
        self.gaussian(self.x, self.t)
        self._project(self.t, self.ph, self.dp, self.a, beta)

    def update(self, beta):
        _lib.call('pgj_gauss_sensitivity', self.n, _lib.ptr(self.dp), _lib.ptr(self.ph), _lib.ptr(self.energy),
                  self.emin, self.penal, _lib.ptr(self.gin), _lib.stream())
        self.gaussian(self.gin, self.g)
        self.gaussian(self.dp, self.dv)
        _lib.call('pgj_gauss_ratio', self.n, _lib.ptr(self.g), _lib.ptr(self.dv), _lib.ptr(self.ratio),
                  _lib.ptr(self.ws), _lib.ptr(self.bounds), _lib.stream())
        for _ in range(self.bisections):
            _lib.call('pgj_gauss_candidate', self.n, _lib.ptr(self.x), _lib.ptr(self.ratio), _lib.ptr(self.bounds),
                      self.move, _lib.ptr(self.new), _lib.stream())
            self.gaussian(self.new, self.t)
            self._project(self.t, self.trial_ph, self.trial_dp, self.trial_a, beta, True)
            _lib.call('pgj_gauss_bisect', self.n, _lib.ptr(self.ws), _lib.ptr(self.bounds), self.volume,
                      _lib.stream())
        self._change()

    def capture(self):
        """The reference captures this loop into graphs (``outer.py:46-54``); here beta is a launch argument of the
        projection, so the loop runs eagerly on the current stream (the same kernels, the same results)."""
        return self

    def evaluate(self, beta):
        self.beta = float(beta)
        self.physical(beta)

    def step(self, energy):
        self.energy.copy_(torch.as_tensor(energy).reshape(self.energy.shape))
        self.update(self.beta)
