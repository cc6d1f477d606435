"""Resident multi-load FP64 PCG with hipGraph-captured iterations.

Mirrors ``exp/asg_gpu_resident_20261103/resident.py:61-213`` (``Resident``)
and the 6-load grouping/energy/C^H of
``exp/asg_sctop_reproduction_20261105/reproduction.py:121-176``
(``PaperResident``): per-load independent recurrences with device-latched
masks, an iteration cap (4000 by default, a per-load device field), up to two
restarts from the certified residual, true-residual certification
``|f - K u| <= max(rtol |f|, 1e-13/sqrt(n))`` and ``CertificationError`` (a
``RuntimeError``, as the reference raises) on failure. Only per-load scalars cross to the host, once
per ``chunk`` iterations.
"""
import math

import numpy as np
import torch

from . import _lib
from ._chunks import MaskPipeline, StreamSplit
from .certify import CAP, MAX_RESTARTS, fail, infos as make_infos, outcome
from .backend import lame
from .operator import LAMBDA_STUDY2, MU_STUDY2, Operator
from .outer import GaussianOuter, PaperOuter

LOADS6 = np.eye(6)


class Resident:
    """Solves L macro-strain loads on one periodic grid, ``batch`` loads per recurrence group.

    Two ways to build it:

    - the reference's, ``Resident(backend, n, chunk=4, load_mode='serial')``
      (``exp/asg_gpu_resident_20261103/resident.py:61-96``): ``backend`` is any
      object with an isotropic 6x6 ``C0`` and ``loads`` (L, 6) (``pgj.backend``);
      the Study-I outer loop (``GaussianOuter``, ``outer.py:12-59``) is attached
      as ``self.outer`` and captured with the solver;
    - ``Resident(n, loads=..., load_mode=..., batch=..., lam=..., mu=...)``:
      the grid size first, everything else by keyword, no outer loop.

    ``load_mode='serial'`` solves one load at a time (batch 1), ``'parallel'``
    groups loads (batch 3 by default, as the reference; any divisor of L may
    be given with ``batch``). ``cap`` is the per-load iteration cap
    (the reference's 4000).
    """

    def __init__(self, backend, n=None, chunk=4, load_mode='serial', *, loads=None, batch=None, lam=None, mu=None,
                 device='cuda', streams=2, cap=CAP, outer=True):
        if hasattr(backend, 'C0') and hasattr(backend, 'loads'):
            if n is None:
                raise TypeError('Resident(backend, n, chunk, load_mode) needs the grid size n')
            blam, bmu = lame(backend.C0)
            lam = blam if lam is None else lam
            mu = bmu if mu is None else mu
            loads = backend.loads if loads is None else loads
            self.b = backend
        else:
            if n is not None:
                raise TypeError('Resident(n, ...) takes the grid size first and the rest by keyword '
                                '(loads=, load_mode=, batch=), or the reference form Resident(backend, n, chunk, '
                                'load_mode)')
            n, self.b, outer = backend, None, False
            lam = LAMBDA_STUDY2 if lam is None else lam
            mu = MU_STUDY2 if mu is None else mu
            loads = LOADS6 if loads is None else loads
        if load_mode not in ('serial', 'parallel'):
            raise ValueError('load_mode must be serial or parallel')
        if chunk < 2 or chunk % 2:
            raise ValueError('graph chunk must be even and >= 2')
        self.loads = np.atleast_2d(np.asarray(loads, dtype=np.float64))
        nl = len(self.loads)
        self.batch = 1 if load_mode == 'serial' else (batch or min(3, nl))
        if nl % self.batch:
            raise ValueError('batch must divide the number of loads')
        if int(cap) < 1:
            raise ValueError('iteration cap must be at least 1')
        self.load_mode, self.chunk, self.cap = load_mode, chunk, int(cap)
        self.n = int(n) if np.isscalar(n) else tuple(int(v) for v in n)
        shape = (n, n, n) if np.isscalar(n) else tuple(n)
        self.shape = shape
        self.op = Operator(torch.ones(shape, dtype=torch.float64, device=device), None, lam, mu, device).enable_green()
        g = self.op.grid
        N = int(np.prod(shape))
        B = self.batch
        dev = torch.device(device)
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)
        # solver vectors are SoA (B, 3, N) on the device; solutions are AoS (L, 3N)
        self.u, self.r, self.p0, self.p1, self.q, self.rhs = (z(B, 3 * N) for _ in range(6))
        self.solutions = z(nl, 3 * N)
        lib = _lib.lib()
        self.state = torch.zeros(lib.pgj_resident_state_bytes(B), dtype=torch.uint8, device=dev)
        offs = (_lib.c_size_t * 7)()
        _lib.check(lib.pgj_resident_state_offsets(B, offs))
        o = list(offs)
        self.rz = self.state[o[0]:o[0] + 8 * B].view(torch.float64)
        self.target2 = self.state[o[1]:o[1] + 8 * B].view(torch.float64)
        self.norm = self.state[o[2]:o[2] + 8 * B].view(torch.float64)
        self.absolute = self.state[o[3]:o[3] + 8 * B].view(torch.float64)
        self.iterations = self.state[o[4]:o[4] + 8 * B].view(torch.int64)
        self.active = self.state[o[5]:o[5] + B]
        self.bad = self.state[o[6]:o[6] + B]
        self._pipeline = MaskPipeline([self.active])
        self._halves = StreamSplit(self.batch, self.batch, streams)
        self.ws = torch.empty(lib.pgj_resident_ws_bytes(g, B), dtype=torch.uint8, device=dev)
        self.scale = torch.ones((1, 3 * N), dtype=torch.float64, device=dev)   # SoA (3, N)
        nz, ny, nx = shape
        self.scale_hat = torch.zeros((3, nz * ny * (nx // 2 + 1)), dtype=torch.complex128, device=dev)
        self._set_scale(0.0)
        self.alpha = torch.zeros((), dtype=torch.float64, device=dev)
        self._selector = None
        self.energy_out = z(nl, N)
        self.ch_ws = torch.empty(lib.pgj_effective_tensor_ws_bytes(g) // 8 + 1, dtype=torch.float64, device=dev)
        self.ch_out = z(36)
        self.init_graph = self.iter_graph = self.cert_graph = None
        self.outer = None
        if outer and isinstance(self.n, int):
            # Study-I outer loop of the reference's Resident (Gaussian filter, tanh projection, sqrt OC)
            self.outer = GaussianOuter(self.n, device=device)

    # -- kernels ---------------------------------------------------------------------
    def to_soa(self, v):
        out = torch.empty_like(v)
        _lib.call('pgj_aos_to_soa', self.op.grid, v.shape[0], _lib.ptr(v), _lib.ptr(out), _lib.stream())
        return out

    def to_aos(self, v):
        out = torch.empty_like(v)
        _lib.call('pgj_soa_to_aos', self.op.grid, v.shape[0], _lib.ptr(v), _lib.ptr(out), _lib.stream())
        return out

    def K(self, u, active=None):
        """Scalar-a operator for AoS (B, 3N) vectors (``resident.py:98``)."""
        return self.op.apply_batched(u, active)

    def true_K(self, u):
        """Certification operator: the independent reference kernel (``resident.py:100-102``)."""
        return self.op.apply_reference(u)

    def B(self, r):
        """P s G s P r for AoS (B, 3N) residuals (``resident.py:104-117``)."""
        s = self.to_aos(self.scale)
        return self.op.green_precond(r, s, self.scale_hat)

    def _set_scale(self, alpha):
        ws = self.op._workspace(1)
        _lib.call('pgj_resident_scale', self.op.grid, _lib.ptr(self.op._plan), _lib.ptr(self.op._material_field),
                  float(alpha), _lib.ptr(self.scale), _lib.ptr(self.scale_hat), _lib.ptr(self.ws), _lib.stream())

    def _setup(self, load_indices, rtol):
        m = self.loads[list(load_indices)].reshape(-1)
        nl = len(load_indices)
        _lib.call('pgj_resident_setup', self.op.grid, nl, nl, 0, nl, _lib.ptr(self.op._material_field),
                  _lib.host_doubles(m), float(rtol), int(self.cap), _lib.ptr(self.rhs), _lib.ptr(self.state), _lib.ptr(self.ws),
                  _lib.stream())

    def certify(self):
        """u <- P u; r = f - K u; absolute = |r| (``resident.py:133-136``)."""
        _lib.call('pgj_resident_certify', self.op.grid, self.batch, self.batch, 0, self.batch,
                  _lib.ptr(self.op._material_field),
                  _lib.ptr(self.u), _lib.ptr(self.rhs), _lib.ptr(self.r), _lib.ptr(self.state), _lib.ptr(self.ws),
                  _lib.stream())

    def initialize(self):
        """active mask and first direction p0 = M^-1 r (``resident.py:119-123``)."""
        _lib.call('pgj_resident_init', self.op.grid, self.batch, self.batch, 0, self.batch, _lib.ptr(self.op._plan),
                  _lib.ptr(self.scale),
                  _lib.ptr(self.scale_hat), _lib.ptr(self.r), _lib.ptr(self.p0), _lib.ptr(self.state),
                  _lib.ptr(self.ws), _lib.stream())

    def _iterate_slice(self, l0, nl, count):
        _lib.call('pgj_resident_iterate', self.op.grid, self.batch, self.batch, l0, nl, _lib.ptr(self.op._plan),
                  _lib.ptr(self.op._material_field), _lib.ptr(self.scale), _lib.ptr(self.scale_hat),
                  _lib.ptr(self.u), _lib.ptr(self.r), _lib.ptr(self.p0), _lib.ptr(self.p1), _lib.ptr(self.q),
                  _lib.ptr(self.state), _lib.ptr(self.ws), count, _lib.stream())

    def iteration(self, count=None):
        """``count`` fused masked PCG iterations (``resident.py:125-131``).

        In the parallel load mode with an even group of four or more loads, the two halves of the group run on
        two streams (forked from and joined to the current one, also inside a captured graph), so the kernels of
        one half fill the tails of the other's. Every load's arithmetic is independent of the other loads, so
        the results are those of one launch over the whole group."""
        count = int(count or self.chunk)
        self._halves.run(lambda l0, nl: self._iterate_slice(l0, nl, count), self.batch)

    def tune(self, reps=6, fft=True):
        """Pick the HEX8 row-march shape and the FFT radix plans by timing them on this solver's buffers at the
        solve's launch shape (one slice of the group per launch); returns the choices (``pgj.tune``). The
        choices hold for every solver of this grid; graphs captured before are captured again."""
        from . import tune as tuner
        parts = self._halves.parts
        nl = parts[0][1] if parts else self.batch
        with tuner.heterogeneous(self.op._material_field):
            out = self._tune(tuner, nl, reps, fft)
        torch.cuda.synchronize()
        self.u.zero_()
        self.solutions.zero_()
        if self.iter_graph is not None:
            self._capture_solver()          # the solver's graphs only: the outer loop's design stays as it is
        return out

    def _tune(self, tuner, nl, reps, fft):
        self._setup(range(self.batch), 1e-8)
        self.r.normal_()
        self.certify()
        self.initialize()

        def stage(k):
            _lib.call('pgj_resident_stage', self.op.grid, self.batch, self.batch, 0, nl, _lib.ptr(self.op._plan),
                      _lib.ptr(self.op._material_field), _lib.ptr(self.scale), _lib.ptr(self.scale_hat),
                      _lib.ptr(self.u), _lib.ptr(self.r), _lib.ptr(self.p0), _lib.ptr(self.p1), _lib.ptr(self.q),
                      _lib.ptr(self.state), _lib.ptr(self.ws), k, _lib.stream())

        for k in range(4):
            stage(k)
        out = {'row_march': tuner.tune_row_march(self.op.grid, nl, 8, lambda: stage(0), reps)}
        if fft and isinstance(self.n, int):
            out['fft'] = tuner.tune_fft(self.op.grid, 8, stage, reps)
        out['active'] = int(self.active.sum().item())     # loads still active at the end of the timing
        return out

    def capture(self):
        """Capture init / chunk-of-iterations / certify graphs (``resident.py:138-150``), and the outer loop's
        (which, as the reference's, warms on a test density: set the outer's initial design after capture)."""
        if self.outer is not None:
            self.outer.capture()
        return self._capture_solver()

    def _capture_solver(self):
        self._setup(range(self.batch), 1e-8)
        self.r.normal_()
        self.certify()
        self.initialize()
        self.iteration()
        torch.cuda.synchronize()
        for name, fn in (('init_graph', self.initialize), ('iter_graph', self.iteration),
                         ('cert_graph', self.certify)):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                fn()
            setattr(self, name, g)
        torch.cuda.synchronize()
        self.u.zero_()
        self.solutions.zero_()
        return self

    # -- design-step API -------------------------------------------------------------
    def material(self, a):
        """Set the SIMP field a (n, n, n) (``resident.py:152-154``)."""
        self.op._material_field.copy_(torch.as_tensor(a, dtype=torch.float64).reshape(self.shape))
        self.op._diagonal = None

    def select(self, method, emin=None, rho=None):
        """alpha by value, 'green' or 'gj', or by a density selector ('eta', 'contrast', 'gray', 'compact8',
        'compact8q') evaluated on the device; rebuilds s and its spectrum (``resident.py:156-174``,
        ``exp/asg_density_selector_20261105/run.py:53-67``). ``emin`` is needed by 'eta', the physical density
        ``rho`` by 'gray'. Returns alpha (a float, or the device scalar for a selector)."""
        from .select import Selector, _METHOD
        if isinstance(method, str) and (method == 'eta' or method in _METHOD) and method not in ('green', 'gj'):
            if self._selector is None:
                self._selector = Selector(self.op.grid, self.op.device)
            alpha = self._selector(self.op._material_field, method, emin, rho)
            self.alpha.copy_(alpha)
            _lib.call('pgj_resident_scale_dev', self.op.grid, _lib.ptr(self.op._plan),
                      _lib.ptr(self.op._material_field), _lib.ptr(self.alpha), _lib.ptr(self.scale),
                      _lib.ptr(self.scale_hat), _lib.ptr(self.ws), _lib.stream())
            return self.alpha
        if method == 'green':
            alpha = 0.0
        elif method == 'gj':
            alpha = 1.0
        else:
            alpha = float(method)
        self.alpha.fill_(alpha)
        self._set_scale(alpha)
        return alpha

    def solve(self, rtol=1e-8):
        """Solve every load; returns (infos, host_checks) (``resident.py:176-184``, ``reproduction.py:144-151``)."""
        infos, checks = [], 0
        for start in range(0, len(self.loads), self.batch):
            idx = list(range(start, start + self.batch))
            _lib.call('pgj_aos_to_soa', self.op.grid, self.batch, _lib.ptr(self.solutions[start:start + self.batch]),
                      _lib.ptr(self.u), _lib.stream())
            info, count = self._solve_batch(idx, rtol)
            _lib.call('pgj_soa_to_aos', self.op.grid, self.batch, _lib.ptr(self.u),
                      _lib.ptr(self.solutions[start:start + self.batch]), _lib.stream())
            infos.extend(info)
            checks += count
        return infos, checks

    def _solve_batch(self, load_indices, rtol):
        """One group: set-up, certify, PCG with restarts, certify (``resident.py:186-209``)."""
        if self.iter_graph is None:
            raise RuntimeError('call capture() first')
        self._setup(load_indices, rtol)
        restarts = host_checks = 0
        self.cert_graph.replay()
        while True:
            self.init_graph.replay()
            host_checks += self._pipeline.run(self.iter_graph.replay)
            self.cert_graph.replay()
            diag = torch.stack((self.norm, self.absolute, self.target2.sqrt(), self.iterations.to(torch.float64),
                                self.bad.to(torch.float64)), dim=1).cpu().numpy()
            ok, stop = outcome(diag, self.cap)
            if ok.all() or restarts >= MAX_RESTARTS or stop:
                break
            restarts += 1
        info = make_infos(diag, ok, restarts)
        if not ok.all():
            fail(info)
        return info, host_checks

    # -- responses ---------------------------------------------------------------------
    def energies(self, solutions=None, loads=None):
        """Element energies (L, N) of the solutions (``energy.cu:7-35``)."""
        sol = self.solutions if solutions is None else solutions
        loads = self.loads if loads is None else np.atleast_2d(loads)
        out = torch.empty((sol.shape[0], sol.shape[1] // 3), dtype=torch.float64, device=sol.device)
        _lib.call('pgj_hex8_energy', self.op.grid, sol.shape[0], _lib.ptr(sol.contiguous()),
                  _lib.host_doubles(np.asarray(loads, np.float64).reshape(-1)), _lib.ptr(out), _lib.stream())
        return out

    def effective_tensor(self):
        """C^H (6, 6) from the six unit-strain solutions (``reproduction.py:166-176``)."""
        if len(self.loads) != 6 or not np.array_equal(self.loads, LOADS6):
            raise ValueError('C^H needs the six unit Voigt strains')
        _lib.call('pgj_effective_tensor', self.op.grid, _lib.ptr(self.op._material_field), _lib.ptr(self.solutions),
                  _lib.ptr(self.ch_ws), _lib.ptr(self.ch_out), _lib.stream())
        return self.ch_out.reshape(6, 6)

    def energy(self):
        """(moduli (L,), sensitivity (n, n, n)): the a-weighted mean energy per load and the load-mean element
        energy (Study I, ``exp/asg_gpu_resident_20261103/resident.py:211-213``)."""
        e = self.energies()
        a = self.op._material_field.reshape(1, -1)
        return (e * a).mean(1), e.mean(0).reshape(self.shape)

    def objective_energy(self, objective='shear', iteration=0):
        """(objective, local sensitivity density, diagonal moduli) of a Study-II objective
        (``reproduction.py:153-164``)."""
        e = self.energies()
        a = self.op._material_field.reshape(1, -1)
        diagonal = (e * a).mean(1)
        if objective == 'shear':
            local = -e[3:].sum(0)
        else:
            hydro_u = self.solutions[:3].sum(0, keepdim=True)
            hydro = self.energies(hydro_u, np.array([[1., 1., 1., 0., 0., 0.]]))[0]
            if objective == 'bulk':
                local = -hydro
            elif objective == 'auxetic':
                local = hydro - (1 + .8 ** iteration) * e[:3].sum(0)
            else:
                raise ValueError(objective)
        obj = (local * a.reshape(-1)).mean()
        return obj, local.reshape(self.shape), diagonal


class PaperResident(Resident):
    """The Study-II resident solver (``exp/asg_sctop_reproduction_20261105/reproduction.py:121-176``).

    ``PaperResident(backend, n, objective, volume, rmin=3., beta=4., eta=.5, emin=1e-9, load_mode='parallel',
    chunk=4)``: all of ``backend.loads`` (the six unit strains) in groups of 3 (parallel) or 1 (serial), the
    cone-filter / eta-projection OC loop (``PaperOuter``) as ``self.outer`` with the reference's move limit 0.1
    unless ``move`` is given, ``energy(iteration) -> (objective, local, diagonal)`` for 'shear', 'bulk' or
    'auxetic', and ``effective_tensor()``.
    """

    def __init__(self, backend, n, objective, volume, rmin=3., beta=4., eta=.5, emin=1e-9, load_mode='parallel',
                 chunk=4, *, move=0.1, batch=None, device='cuda', streams=2, cap=CAP):
        if load_mode not in ('serial', 'parallel'):
            raise ValueError(load_mode)
        if objective not in ('shear', 'bulk', 'auxetic'):
            raise ValueError(objective)
        super().__init__(backend, n, chunk, load_mode, batch=batch, device=device, streams=streams, cap=cap,
                         outer=False)
        self.objective_kind = objective
        self.outer = PaperOuter(int(n), volume, rmin=rmin, beta=beta, eta=eta, emin=emin, move=move, device=device)

    def energy(self, iteration=0):
        """(objective, local (n, n, n), diagonal (6,)) (``reproduction.py:153-164``)."""
        return self.objective_energy(self.objective_kind, iteration)

    def select(self, method, emin=None, rho=None):
        """``Resident.select`` where the 'eta' table reads the outer loop's Emin and the 'gray' selector its
        physical density, as the reference's selector resident does
        (``exp/asg_density_selector_20261105/run.py:46-77``)."""
        return super().select(method, self.outer.emin if emin is None else emin,
                              self.outer.ph if rho is None else rho)


# The density-selector resident of exp/asg_density_selector_20261105/run.py:46 (PaperResident with the
# selectors; here every Resident selects on the device).
SelectorResident = PaperResident
