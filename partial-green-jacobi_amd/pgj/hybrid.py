"""General-C21 (multi-phase, anisotropic) operator on libpgj's hybrid kernels.

``HybridFIOperator`` keeps the constructor, attributes, methods and errors of
``solver/src/materials/hybrid_adapter.py:39-125``: element e has the stiffness
of pool row ``index[e]`` (``dense=False``) or row e (``dense=True``), built by
``pgj.materials.build_hybrid_material`` / ``build_material_field``.
``enable_fi_reference`` attaches the FI Green preconditioner of the isotropic
'zeman' reference medium (``solver/src/reference_medium/formula_adapter.py:40-68``),
so the operator plugs into ``solve_pcg_batched`` and
``compute_effective_stiffness_batched``. Beyond the reference, ``apply_batched``
applies K to (L, 3N) loads in one launch.
"""
from types import MethodType, SimpleNamespace

import numpy as np
import torch

from . import _lib
from .materials import PAIRS


class HybridFIOperator:
    """B=1 periodic FI operator whose element stiffnesses come from a C21 pool."""

    def __init__(self, coefficients, index, grid, voxel, *, dense=False):
        nz, ny, nx = map(int, grid)
        if min(nz, ny, nx) < 2 or min(voxel) <= 0:
            raise ValueError('Periodic grid dimensions >=2 and positive voxel sizes required')
        self.B = 1
        self.device, self.dtype = coefficients.device, coefficients.dtype
        if self.device.type != 'cuda' or self.dtype != torch.float64:
            raise ValueError('FI hybrid operator requires CUDA FP64')
        n = nx * ny * nz
        self.mesh = SimpleNamespace(Nz=nz, Ny=ny, Nx=nx, dx=float(voxel[0]), dy=float(voxel[1]),
                                    dz=float(voxel[2]), n_elem=n, n_nodes=n, n_dofs=3 * n,
                                    V_elem=float(np.prod(voxel)))
        coefficients = coefficients.reshape(-1, 21)
        if not coefficients.is_contiguous():
            raise ValueError('C21 coefficients must be contiguous')
        self._hybrid_coeff = coefficients
        self._hybrid_index = index
        self._hybrid_indirect = not dense
        rows = coefficients.shape[0]
        if dense:
            if rows != n:
                raise ValueError(f'dense storage needs {n} C21 rows, got {rows}')
            self._index_ptr = None
        else:
            if index.dtype != torch.int32 or index.device != self.device or not index.is_contiguous() \
                    or index.numel() != n:
                raise ValueError('int32 contiguous row map of N entries on the material device required')
            lo, hi = int(index.min()), int(index.max())
            if lo < 0 or hi >= rows:
                raise ValueError(f'row map entries must lie in [0, {rows}), got [{lo}, {hi}]')
            self._index_ptr = index
        self.C = None
        self._C_voigt_gpu = None
        self.element_type = 'c3d8'
        self._modal_kind = 'fi'
        self._modal_voxel = tuple(map(float, voxel))
        self._fft_enabled = False
        self._diag_cache = self._M_inv_cache = None
        self._fft_symbol_inv = None
        self._fft_args = None
        self.grid = _lib.make_grid((nz, ny, nx), self._modal_voxel, 0.0, 0.0)
        self._stress_ws = torch.empty(_lib.lib().pgj_hybrid_stress_ws_bytes(self.grid) // 8 + 1,
                                      dtype=torch.float64, device=self.device)
        self._proj_ws = None

    def geometry(self):
        m = self.mesh
        return (m.Nz, m.Ny, m.Nx, m.dx, m.dy, m.dz)

    def _validate_apply_input(self, u):
        if u.shape != (1, self.mesh.n_dofs) or u.dtype != self.dtype or u.device != self.device \
                or not u.is_contiguous():
            raise ValueError('Expected contiguous FP64 displacement (1,3N) on the material device')

    def project_zero_mean(self, value):
        """Subtract the per-component nodal mean (``hybrid_adapter.py:78-80``); returns a new (1, 3N) tensor."""
        out = value.reshape(1, -1).contiguous().clone()
        need = _lib.lib().pgj_project_ws_bytes(self.grid, 1) // 8 + 1
        if self._proj_ws is None or self._proj_ws.numel() < need:
            self._proj_ws = torch.empty(need, dtype=torch.float64, device=self.device)
        _lib.call('pgj_project', self.grid, 1, _lib.ptr(out), _lib.ptr(self._proj_ws), _lib.stream())
        return out

    def _macro_strain_tensor(self, macro):
        m = macro.detach().cpu().numpy() if torch.is_tensor(macro) else macro
        if np.shape(m) != (6,):
            raise ValueError('Six engineering macro strain components required')
        return _lib.host_doubles(np.asarray(m, dtype=np.float64))

    def _pool(self):
        return _lib.ptr(self._hybrid_coeff), _lib.ptr(self._index_ptr)

    def apply(self, u):
        """K u for one (1, 3N) displacement (``hybrid_adapter.py:87-90``)."""
        self._validate_apply_input(u)
        out = torch.empty_like(u)
        _lib.call('pgj_hybrid_apply', self.grid, 1, *self._pool(), _lib.ptr(u), None, _lib.ptr(out), _lib.stream())
        return out

    def apply_batched(self, u, active=None):
        """K u_l for (L, 3N) loads in one launch; loads with ``active[l] == 0`` give zeros."""
        if u.dim() != 2 or u.shape[1] != self.mesh.n_dofs or u.dtype != self.dtype or u.device != self.device \
                or not u.is_contiguous():
            raise ValueError('Expected contiguous FP64 displacements (L,3N) on the material device')
        if active is not None:
            active = active.to(device=self.device, dtype=torch.uint8).contiguous()
            if active.numel() != u.shape[0]:
                raise ValueError('one active flag per load required')
        out = torch.empty_like(u)
        _lib.call('pgj_hybrid_apply', self.grid, u.shape[0], *self._pool(), _lib.ptr(u), _lib.ptr(active),
                  _lib.ptr(out), _lib.stream())
        return out

    def apply_gather(self, u, active=None):
        """The same K u_l by the independent thread-per-node gather kernel (certification path)."""
        if u.dim() != 2 or u.shape[1] != self.mesh.n_dofs or u.dtype != self.dtype or u.device != self.device \
                or not u.is_contiguous():
            raise ValueError('Expected contiguous FP64 displacements (L,3N) on the material device')
        if active is not None:
            active = active.to(device=self.device, dtype=torch.uint8).contiguous()
        out = torch.empty_like(u)
        _lib.call('pgj_hybrid_apply_gather', self.grid, u.shape[0], *self._pool(), _lib.ptr(u), _lib.ptr(active),
                  _lib.ptr(out), _lib.stream())
        return out

    def compute_rhs(self, macro):
        """f = -V sum_e Bbar^T C_e eps(macro), (1, 3N) (``hybrid_adapter.py:92-94``)."""
        m = self._macro_strain_tensor(macro)
        out = torch.empty((1, self.mesh.n_dofs), dtype=self.dtype, device=self.device)
        _lib.call('pgj_hybrid_rhs', self.grid, *self._pool(), m, _lib.ptr(out), _lib.stream())
        return out

    def compute_avg_stress(self, u, macro):
        """Mean stress <C (eps(macro) + Bbar u)> as numpy (1, 6) (``hybrid_adapter.py:96-99``)."""
        self._validate_apply_input(u)
        m = self._macro_strain_tensor(macro)
        out = torch.empty(6, dtype=self.dtype, device=self.device)
        _lib.call('pgj_hybrid_avg_stress', self.grid, *self._pool(), _lib.ptr(u), m, _lib.ptr(self._stress_ws),
                  _lib.ptr(out), _lib.stream())
        return out.cpu().numpy().reshape(1, 6)

    def _compute_reference_medium(self, strategy, *, verbose=False):
        """Isotropic 'zeman' reference medium from a 16-per-axis sample of element rows
        (``hybrid_adapter.py:101-125``): mu, K midway between the sampled extremes of
        (C33+C44+C55)/6 and (C00+C11+C22+2(C01+C02+C12))/9, lam = max(0, K - 2mu/3)."""
        if strategy != 'zeman':
            raise ValueError('This operator uses the zeman reference medium only')
        m = self.mesh
        z = np.arange(0, m.Nz, max(1, m.Nz // 16))
        y = np.arange(0, m.Ny, max(1, m.Ny // 16))
        x = np.arange(0, m.Nx, max(1, m.Nx // 16))
        positions = ((z[:, None, None] * m.Ny + y[None, :, None]) * m.Nx + x[None, None, :]).reshape(-1)
        pos = torch.as_tensor(positions, device=self.device)
        rows = self._hybrid_index.reshape(-1)[pos].long() if self._hybrid_indirect else pos
        packed = self._hybrid_coeff.index_select(0, rows).cpu().numpy()
        C = np.zeros((len(packed), 6, 6))
        for k, (i, j) in enumerate(PAIRS):
            C[:, i, j] = C[:, j, i] = packed[:, k]
        mus = 0.5 * (C[:, 3, 3] + C[:, 4, 4] + C[:, 5, 5]) / 2.0
        Ks = (C[:, 0, 0] + C[:, 1, 1] + C[:, 2, 2] + 2 * (C[:, 0, 1] + C[:, 0, 2] + C[:, 1, 2])) / 9.0
        mu = 0.5 * (float(mus.max()) + float(mus.min()))
        K = 0.5 * (float(Ks.max()) + float(Ks.min()))
        lam = max(0.0, K - 2 * mu / 3.0)
        out = np.zeros((6, 6))
        out[:3, :3] = lam
        out[np.arange(3), np.arange(3)] += 2 * mu
        out[np.arange(3, 6), np.arange(3, 6)] = mu
        return out[None]


# cubes with compile-time FFT passes, the grids of the FP32 apply
FP32_CUBES = (16, 24, 32, 48, 64, 96, 128)


def enable_fi_reference(op, precision='mixed', layout='upper6'):
    """Attach ``apply_precond`` = the FI Green operator of the isotropic reference medium
    (``formula_adapter.py:40-68``). ``precision='mixed'`` (the default) applies it in FP32 as the reference
    does (``formula_adapter.py:59-66``: the residual rounded to FP32, FP32 transforms and symbol, the result
    returned in the residual's dtype) on the cubes in ``FP32_CUBES``, and in FP64 on other grids;
    ``'fp64'`` applies it in FP64 everywhere."""
    if precision not in ('mixed', 'fp64'):
        raise ValueError(precision)
    reference = op._compute_reference_medium('zeman', verbose=False)[0]
    lam, mu = float(reference[0, 1]), float(reference[3, 3])
    expected = np.zeros((6, 6))
    expected[:3, :3] = lam
    expected[np.arange(3), np.arange(3)] += 2 * mu
    expected[np.arange(3, 6), np.arange(3, 6)] = mu
    if not np.allclose(expected, reference, rtol=1e-14, atol=0):
        raise ValueError('Isotropic reference required')
    shape = (op.mesh.Nz, op.mesh.Ny, op.mesh.Nx)
    op.grid = _lib.make_grid(shape, (op.mesh.dx, op.mesh.dy, op.mesh.dz), lam, mu)
    lib = _lib.lib()
    op._plan = torch.empty(lib.pgj_plan_bytes(op.grid), dtype=torch.uint8, device=op.device)
    _lib.call('pgj_plan_init', op.grid, _lib.ptr(op._plan), _lib.stream())
    op._pre_ws = torch.empty(lib.pgj_precond_ws_bytes(op.grid, 1), dtype=torch.uint8, device=op.device)
    op._plan32 = None
    if precision == 'mixed' and shape[0] == shape[1] == shape[2] and shape[0] in FP32_CUBES:
        op._plan32 = torch.empty(lib.pgj_plan32_bytes(op.grid), dtype=torch.uint8, device=op.device)
        _lib.call('pgj_plan32_init', op.grid, _lib.ptr(op._plan), _lib.ptr(op._plan32), _lib.stream())
    op._formula_reference_C0 = reference.tolist()
    op._fft_enabled = True
    op._fft_spectrum_mode = 'real_half'
    op._fft_symbol_layout = 'FI_diagonal_rank1'
    op._fft_inverse_norm_folded = True
    op._fft_batch_chunk_size = None
    op._M_inv_cache = None
    op._formula_precision = precision

    def apply(self, residual):
        r = residual.reshape(1, -1).contiguous()
        if self._plan32 is not None:
            z = torch.empty_like(r)
            _lib.call('pgj_precond_apply_mixed', self.grid, 1, _lib.ptr(self._plan32), _lib.ptr(r), None, None,
                      _lib.ptr(z), _lib.ptr(self._pre_ws), _lib.stream())
            return z.view_as(residual)
        z = torch.empty_like(r)
        _lib.call('pgj_precond_apply', self.grid, 1, _lib.ptr(self._plan), _lib.ptr(r), None, None, _lib.ptr(z),
                  _lib.ptr(self._pre_ws), _lib.stream())
        return z.view_as(residual)

    op.apply_precond = MethodType(apply, op)
    return op


class HybridResident:
    """The resident multi-load PCG over a general-C21 operator: ``op`` is a ``HybridFIOperator`` with
    ``enable_fi_reference`` applied (its grid carries the Green reference medium and its plan).

    The recurrence, masks, iteration cap, restarts and true-residual certification are those of
    ``pgj.Resident`` (``exp/asg_gpu_resident_20261103/resident.py:39-58``, ``:186-209``). The preconditioner is
    the operator's unscaled FI Green operator, as ``apply_precond``, applied in FP64 by the fused stages whatever
    ``enable_fi_reference``'s ``precision`` (its FP32 apply serves the generic path). The loads of a group
    advance together in hipGraph-captured chunks of ``chunk`` iterations:
    - ``q = K p`` is the tiled C21 apply with fused p.q partials;
    - the three FFT passes are those of the scalar solver;
    - certification uses the gather apply.
    In the parallel load mode (the default, one group of all loads) the group runs as ``streams`` slices on as
    many streams. Three 2-load slices measured 150.5 solves/s at the 64^3 composite, against 139.8 for two
    halves and 145.2 for one stream.

    This is the resident counterpart of ``compute_effective_stiffness_batched`` over this operator
    (``solver/src/runtime/fi/source/femlib/batched/driver.py:13-75``), which solves one load at a time with
    the generic PCG. ``effective_tensor`` gives C^H from the mean stresses of the six solutions.
    """

    def __init__(self, op, loads=None, load_mode='parallel', batch=None, chunk=4, streams=3, cap=None):
        from .certify import CAP
        from ._chunks import MaskPipeline, StreamSplit
        if not getattr(op, '_fft_enabled', False) or getattr(op, '_plan', None) is None:
            raise ValueError('enable_fi_reference(op) first: the solver uses its Green plan')
        if load_mode not in ('serial', 'parallel'):
            raise ValueError('load_mode must be serial or parallel')
        if chunk < 2 or chunk % 2:
            raise ValueError('graph chunk must be even and >= 2')
        self.op = op
        self.loads = np.eye(6) if loads is None else np.atleast_2d(np.asarray(loads, dtype=np.float64))
        nl = len(self.loads)
        self.batch = 1 if load_mode == 'serial' else (batch or nl)
        if nl % self.batch:
            raise ValueError('batch must divide the number of loads')
        self.cap = CAP if cap is None else int(cap)
        if self.cap < 1:
            raise ValueError('iteration cap must be at least 1')
        self.load_mode, self.chunk = load_mode, chunk
        m = op.mesh
        N, B, dev = m.n_nodes, self.batch, op.device
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device=dev)
        self.u, self.r, self.p0, self.p1, self.q, self.rhs = (z(B, 3 * N) for _ in range(6))
        self.solutions = z(nl, 3 * N)
        lib = _lib.lib()
        g = op.grid
        self.state = torch.zeros(lib.pgj_resident_state_bytes(B), dtype=torch.uint8, device=dev)
        offs = (_lib.c_size_t * 7)()
        _lib.check(lib.pgj_resident_state_offsets(B, offs))
        o = list(offs)
        self.target2 = self.state[o[1]:o[1] + 8 * B].view(torch.float64)
        self.norm = self.state[o[2]:o[2] + 8 * B].view(torch.float64)
        self.absolute = self.state[o[3]:o[3] + 8 * B].view(torch.float64)
        self.iterations = self.state[o[4]:o[4] + 8 * B].view(torch.int64)
        self.active = self.state[o[5]:o[5] + B]
        self.bad = self.state[o[6]:o[6] + B]
        self.ws = torch.empty(lib.pgj_resident_ws_bytes(g, B), dtype=torch.uint8, device=dev)
        self._pipeline = MaskPipeline([self.active])
        self._halves = StreamSplit(B, B, streams)
        self._pool_args = op._pool()
        self.init_graph = self.iter_graph = self.cert_graph = None

    def _soa(self, src, dst, L):
        _lib.call('pgj_aos_to_soa', self.op.grid, L, _lib.ptr(src), _lib.ptr(dst), _lib.stream())

    def _aos(self, src, dst, L):
        _lib.call('pgj_soa_to_aos', self.op.grid, L, _lib.ptr(src), _lib.ptr(dst), _lib.stream())

    def _setup(self, load_indices, rtol):
        m = self.loads[list(load_indices)].reshape(-1)
        nl = len(load_indices)
        _lib.call('pgj_hybrid_resident_setup', self.op.grid, nl, 0, nl, *self._pool_args, _lib.host_doubles(m),
                  float(rtol), int(self.cap), _lib.ptr(self.rhs), _lib.ptr(self.state), _lib.ptr(self.ws),
                  _lib.stream())

    def certify(self):
        """u <- P u; r = f - K u (gather apply); absolute = |r|."""
        _lib.call('pgj_hybrid_resident_certify', self.op.grid, self.batch, 0, self.batch, *self._pool_args,
                  _lib.ptr(self.u), _lib.ptr(self.rhs), _lib.ptr(self.r), _lib.ptr(self.state), _lib.ptr(self.ws),
                  _lib.stream())

    def initialize(self):
        """active mask and first direction p0 = G r."""
        _lib.call('pgj_resident_init', self.op.grid, self.batch, self.batch, 0, self.batch, _lib.ptr(self.op._plan),
                  None, None, _lib.ptr(self.r), _lib.ptr(self.p0), _lib.ptr(self.state), _lib.ptr(self.ws),
                  _lib.stream())

    def _iterate_slice(self, l0, nl, count):
        _lib.call('pgj_hybrid_resident_iterate', self.op.grid, self.batch, l0, nl, _lib.ptr(self.op._plan),
                  *self._pool_args, None, None, _lib.ptr(self.u), _lib.ptr(self.r), _lib.ptr(self.p0),
                  _lib.ptr(self.p1), _lib.ptr(self.q), _lib.ptr(self.state), _lib.ptr(self.ws), count,
                  _lib.stream())

    def iteration(self, count=None):
        """``count`` masked PCG iterations of the group (its halves on two streams in the parallel mode)."""
        count = int(count or self.chunk)
        self._halves.run(lambda l0, nl: self._iterate_slice(l0, nl, count), self.batch)

    def capture(self):
        """Capture the init / chunk / certify graphs (``resident.py:138-150``)."""
        self._setup(range(self.batch), 1e-8)
        self.r.normal_()
        self.certify()
        self.initialize()
        self.iteration()
        torch.cuda.synchronize()
        for name, fn in (('init_graph', self.initialize), ('iter_graph', self.iteration),
                         ('cert_graph', self.certify)):
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                fn()
            setattr(self, name, graph)
        torch.cuda.synchronize()
        self.u.zero_()
        self.solutions.zero_()
        return self

    def solve(self, rtol=1e-8):
        """Solve every load from the current solutions (warm start); returns (infos, host_checks) with the
        resident solver's info keys. Raises ``CertificationError`` if a load cannot be certified."""
        from .certify import MAX_RESTARTS, fail, infos as make_infos, outcome
        if self.iter_graph is None:
            raise RuntimeError('call capture() first')
        infos, checks = [], 0
        for start in range(0, len(self.loads), self.batch):
            idx = list(range(start, start + self.batch))
            sol = self.solutions[start:start + self.batch]
            self._soa(sol, self.u, self.batch)
            self._setup(idx, rtol)
            restarts = 0
            self.cert_graph.replay()
            while True:
                self.init_graph.replay()
                checks += self._pipeline.run(self.iter_graph.replay)
                self.cert_graph.replay()
                diag = torch.stack((self.norm, self.absolute, self.target2.sqrt(),
                                    self.iterations.to(torch.float64), self.bad.to(torch.float64)),
                                   dim=1).cpu().numpy()
                ok, stop = outcome(diag, self.cap)
                if ok.all() or restarts >= MAX_RESTARTS or stop:
                    break
                restarts += 1
            info = make_infos(diag, ok, restarts)
            self._aos(self.u, sol, self.batch)
            if not ok.all():
                fail(info)
            infos.extend(info)
        return infos, checks

    def effective_tensor(self):
        """C^H (6, 6): column j is the mean stress of solution j under unit strain e_j."""
        if len(self.loads) != 6 or not np.array_equal(self.loads, np.eye(6)):
            raise ValueError('C^H needs the six unit Voigt strains')
        cols = [self.op.compute_avg_stress(self.solutions[j:j + 1].contiguous(), np.eye(6)[j])[0] for j in range(6)]
        return np.stack(cols, axis=1)


def compute_effective_stiffness_resident(op, tol=1e-8, *, solver=None):
    """C^H (1, 6, 6) of a general-C21 operator through ``HybridResident`` (captured on first use; pass a solver
    to reuse its graphs and warm starts), with the per-load infos: the resident counterpart of
    ``compute_effective_stiffness_batched`` (``solver/src/runtime/fi/source/femlib/batched/driver.py:13-75``)."""
    solver = HybridResident(op).capture() if solver is None else solver
    infos, _ = solver.solve(tol)
    C = torch.tensor(solver.effective_tensor()[None], dtype=torch.float64)
    return C, infos
