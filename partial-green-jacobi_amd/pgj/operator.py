"""Drop-in operator and partial Green-Jacobi preconditioner on libpgj.

``Operator`` exposes the duck-typed surface of the reference's
``HybridFIOperator`` (``solver/src/materials/hybrid_adapter.py:39-125``) for a
scalar SIMP field a(x) times an isotropic C0: ``B``, ``device``, ``dtype``,
``mesh``, ``geometry()``, ``apply``, ``compute_rhs``, ``compute_avg_stress``,
``project_zero_mean``, plus what ``build`` attaches
(``solver/experiments/adaptive_green_scaling/study.py:61-77``):
``apply_precond``, ``_unscaled_green``, ``_material_field``, ``_voxel``,
``_diagonal``, ``_log_diagonal``, ``_contest_coefficient`` (the 6-coefficient
symbol) and ``_contest_metadata``. ``diagonal`` and ``set_alpha`` follow
``adaptive_green_scaling/study.py:80-101``.
"""
from types import SimpleNamespace

import numpy as np
import torch

from . import _lib

LAMBDA_STUDY2, MU_STUDY2 = 15 / 26, 5 / 13   # E = 1, nu = 0.3 (reproduction.py:59-62)


def _f64(x, device):
    return torch.as_tensor(x, dtype=torch.float64, device=device).contiguous()


class Operator:
    """B=1 periodic Q1/HEX8 operator for C(x) = a(x) C0(lam, mu) on a CUDA device.

    ``dtype=torch.float32`` gives the inner operator of iterative refinement
    (``solver/src/runtime/fi/source/femlib/batched/ir.py:33-35``): ``apply``
    runs the FP32 row-march on an FP32 copy of a; the preconditioner, rhs,
    stress and projection compute in FP64 and return the operator's dtype.
    """

    def __init__(self, a, voxel=None, lam=LAMBDA_STUDY2, mu=MU_STUDY2, device='cuda', dtype=torch.float64):
        a = _f64(a, device)
        if a.dim() != 3 or min(a.shape) < 2:
            raise ValueError('Periodic grid dimensions >=2 required')
        nz, ny, nx = a.shape
        voxel = tuple(1.0 / np.array((nx, ny, nz))) if voxel is None else tuple(map(float, voxel))
        if min(voxel) <= 0:
            raise ValueError('positive voxel sizes required')
        self.B = 1
        if dtype not in (torch.float64, torch.float32):
            raise ValueError('operator dtype must be float64 or float32')
        self.device, self.dtype = a.device, dtype
        if self.device.type != 'cuda':
            raise ValueError('pgj operator requires a CUDA (HIP) device')
        n = nx * ny * nz
        self.mesh = SimpleNamespace(Nz=nz, Ny=ny, Nx=nx, dx=voxel[0], dy=voxel[1], dz=voxel[2], n_elem=n,
                                    n_nodes=n, n_dofs=3 * n, V_elem=float(np.prod(voxel)))
        self.lam, self.mu = float(lam), float(mu)
        self.grid = _lib.make_grid(a.shape, voxel, lam, mu)
        self._material_field = a
        self._material32 = a.float().contiguous() if dtype == torch.float32 else None
        self._voxel = voxel
        self._diagonal = None
        self._log_diagonal = None
        self._proj_ws = torch.empty(_lib.lib().pgj_project_ws_bytes(self.grid, 1) // 8 + 1, dtype=torch.float64,
                                    device=self.device)
        self._stress_ws = torch.empty(_lib.lib().pgj_avg_stress_ws_bytes(self.grid) // 8 + 1, dtype=torch.float64,
                                      device=self.device)

    def geometry(self):
        m = self.mesh
        return (m.Nz, m.Ny, m.Nx, m.dx, m.dy, m.dz)

    def _validate_apply_input(self, u):
        if u.shape != (1, self.mesh.n_dofs) or u.dtype != self.dtype or u.device != self.device \
                or not u.is_contiguous():
            name = 'FP64' if self.dtype == torch.float64 else 'FP32'
            raise ValueError(f'Expected contiguous {name} displacement (1,3N) on the material device')

    def _apply_call(self, L, u, active, out):
        if self.dtype == torch.float32:
            _lib.call('pgj_hex8_apply32', self.grid, L, _lib.ptr(self._material32), _lib.ptr(u), _lib.ptr(active),
                      _lib.ptr(out), _lib.stream())
        else:
            _lib.call('pgj_hex8_apply', self.grid, L, _lib.ptr(self._material_field), _lib.ptr(u), _lib.ptr(active),
                      _lib.ptr(out), _lib.stream())

    def apply(self, u):
        """K(a) u for one (1, 3N) displacement (``hybrid_adapter.py:87-90``)."""
        self._validate_apply_input(u)
        out = torch.empty_like(u)
        self._apply_call(1, u, None, out)
        return out

    def apply_batched(self, u, active=None):
        """K(a) u_l for (L, 3N) loads; inactive loads give zeros (``scalar_fi.cu:46-53``)."""
        if u.dim() != 2 or u.shape[1] != self.mesh.n_dofs or u.dtype != self.dtype or not u.is_contiguous():
            raise ValueError('Expected contiguous displacements (L,3N) of the operator dtype')
        out = torch.empty_like(u)
        if active is not None:
            active = active.to(torch.uint8).contiguous()
        self._apply_call(u.shape[0], u, active, out)
        return out

    def apply_batched_mfma(self, u, active=None):
        """``apply_batched`` of the FP32 operator with the element product as a dense 24x24 Ke contraction on
        the matrix cores (``pgj_hex8_apply32_mfma``) instead of the modal row-march."""
        if self.dtype != torch.float32:
            raise ValueError('the matrix-core apply is the FP32 operator')
        if u.dim() != 2 or u.shape[1] != self.mesh.n_dofs or u.dtype != self.dtype or not u.is_contiguous():
            raise ValueError('Expected contiguous displacements (L,3N) of the operator dtype')
        out = torch.empty_like(u)
        if active is not None:
            active = active.to(torch.uint8).contiguous()
        _lib.call('pgj_hex8_apply32_mfma', self.grid, u.shape[0], _lib.ptr(self._material32), _lib.ptr(u),
                  _lib.ptr(active), _lib.ptr(out), 0, _lib.stream())
        return out

    def apply_reference(self, u):
        """Same operator by the independent thread-per-node kernel (certification operator)."""
        out = torch.empty_like(u)
        _lib.call('pgj_hex8_apply_reference', self.grid, u.shape[0], _lib.ptr(self._material_field), _lib.ptr(u),
                  _lib.ptr(out), _lib.stream())
        return out

    def _macro(self, macro):
        if np.shape(macro) != (6,):
            raise ValueError('Six engineering macro strain components required')
        return _lib.host_doubles(np.asarray(macro, dtype=np.float64))

    def compute_rhs(self, macro):
        """f = -sum_e a_e Ke chi(macro), (1, 3N) (``hybrid_adapter.py:92-94``)."""
        m = macro.tolist() if torch.is_tensor(macro) else macro
        out = torch.empty((1, self.mesh.n_dofs), dtype=torch.float64, device=self.device)
        _lib.call('pgj_hex8_rhs', self.grid, _lib.ptr(self._material_field), self._macro(m), _lib.ptr(out),
                  _lib.stream())
        return out.to(self.dtype)

    def compute_avg_stress(self, u, macro):
        """Mean stress <C(eps + B u)> as numpy (1, 6) (``hybrid_adapter.py:96-99``)."""
        self._validate_apply_input(u)
        m = macro.tolist() if torch.is_tensor(macro) else macro
        u = u.to(torch.float64)
        out = torch.empty(6, dtype=torch.float64, device=self.device)
        _lib.call('pgj_hex8_avg_stress', self.grid, _lib.ptr(self._material_field), _lib.ptr(u), self._macro(m),
                  _lib.ptr(self._stress_ws), _lib.ptr(out), _lib.stream())
        return out.cpu().numpy().reshape(1, 6)

    def project_zero_mean(self, value):
        """Zero-mean projection per component (``hybrid_adapter.py:78-80``); returns a new tensor."""
        out = value.reshape(value.shape[0], -1).to(torch.float64).contiguous().clone()
        L = out.shape[0]
        ws = self._proj_ws
        need = _lib.lib().pgj_project_ws_bytes(self.grid, L) // 8 + 1
        if ws.numel() < need:
            ws = self._proj_ws = torch.empty(need, dtype=torch.float64, device=self.device)
        _lib.call('pgj_project', self.grid, L, _lib.ptr(out), _lib.ptr(ws), _lib.stream())
        return out.reshape(value.shape).to(value.dtype)

    # -- spectral Green part -------------------------------------------------------------
    def enable_green(self):
        """Plan (twiddles + dense FI symbol) (``formula_adapter.py:40-68``, ``contest_adapter.py:182-230``)."""
        lib = _lib.lib()
        self._plan = torch.empty(lib.pgj_plan_bytes(self.grid), dtype=torch.uint8, device=self.device)
        _lib.call('pgj_plan_init', self.grid, _lib.ptr(self._plan), _lib.stream())
        self._pre_ws = {}
        nz, ny, nx = self.mesh.Nz, self.mesh.Ny, self.mesh.Nx
        self._contest_metadata = dict(kind='fi', variant='dense', grid=(nz, ny, nx), voxel=self._voxel,
                                      physical_precision='FP64', FFT_precision='FP64/complex128',
                                      normalization='folded once into coefficients', native='libpgj LDS Stockham')
        return self

    def _fi_tables(self):
        if getattr(self, '_fi_cache', None) is None:
            nz, ny, nx = self.mesh.Nz, self.mesh.Ny, self.mesh.Nx
            nkx = nx // 2 + 1
            out4 = torch.empty((4, nz * ny * nkx), dtype=torch.float64, device=self.device)
            vtab = torch.empty(nkx + ny + nz, dtype=torch.float64, device=self.device)
            _lib.call('pgj_fi_coeff', self.grid, _lib.ptr(out4), _lib.ptr(vtab), _lib.stream())
            self._fi_cache = (out4, [vtab[:nkx], vtab[nkx:nkx + ny], vtab[nkx + ny:]])
        return self._fi_cache

    @property
    def _fft_symbol_inv(self):
        """(4, nf) factored symbol D_00, D_11, D_22, beta (``formula_adapter.py:54``)."""
        return self._fi_tables()[0]

    @property
    def _formula_vtables(self):
        """Per-axis v tables [x (nx/2+1), y (ny), z (nz)] (``formula_adapter.py:49-53``)."""
        return self._fi_tables()[1]

    @property
    def _contest_plan(self):
        """The plan buffer (twiddles and the dense symbol) the Green apply reads."""
        return self._plan

    @property
    def _hybrid_coeff(self):
        """(N, 21) general-C21 rows a_e C0 of this scalar field (``hybrid_adapter.py:57``), built on demand."""
        c0 = np.zeros((6, 6))
        c0[:3, :3] = self.lam
        c0[np.arange(3), np.arange(3)] += 2 * self.mu
        c0[np.arange(3, 6), np.arange(3, 6)] = self.mu
        row = torch.tensor(c0[np.triu_indices(6)], dtype=torch.float64, device=self.device)
        return self._material_field.reshape(-1, 1) * row[None]

    @property
    def _hybrid_index(self):
        return torch.empty(0, dtype=torch.int32, device=self.device)

    _hybrid_indirect = False

    @property
    def _contest_coefficient(self):
        nz, ny, nx = self.mesh.Nz, self.mesh.Ny, self.mesh.Nx
        out = torch.empty((6, nz * ny * (nx // 2 + 1)), dtype=torch.float64, device=self.device)
        _lib.call('pgj_plan_symbol', self.grid, _lib.ptr(self._plan), _lib.ptr(out), _lib.stream())
        return out

    def _workspace(self, L):
        ws = self._pre_ws.get(L)
        if ws is None:
            ws = self._pre_ws[L] = torch.empty(_lib.lib().pgj_precond_ws_bytes(self.grid, L), dtype=torch.uint8,
                                               device=self.device)
        return ws

    def scale_spectrum(self, s):
        """Spectrum of the (1, 3N) scale s, used by the fused projection of M_alpha."""
        nz, ny, nx = self.mesh.Nz, self.mesh.Ny, self.mesh.Nx
        s = s.reshape(1, -1).contiguous()
        shat = torch.empty((3, nz * ny * (nx // 2 + 1)), dtype=torch.complex128, device=self.device)
        _lib.call('pgj_scale_spectrum', self.grid, _lib.ptr(self._plan), _lib.ptr(s), _lib.ptr(shat),
                  _lib.ptr(self._workspace(1)), _lib.stream())
        return shat

    def green_precond(self, r, s=None, shat=None):
        """z = P s G s P r for (L, 3N) residuals (s None: P G P r), in FP64; FP32 residuals come back FP32."""
        dt = r.dtype
        r = r.reshape(r.shape[0], -1).to(torch.float64).contiguous()
        z = torch.empty_like(r)
        _lib.call('pgj_precond_apply', self.grid, r.shape[0], _lib.ptr(self._plan), _lib.ptr(r),
                  _lib.ptr(s), _lib.ptr(shat), _lib.ptr(z), _lib.ptr(self._workspace(r.shape[0])), _lib.stream())
        return z.to(dt)

    def _unscaled_green(self, r):
        return self.green_precond(r)


def build(rho, lam=LAMBDA_STUDY2, mu=MU_STUDY2, device='cuda', dtype=torch.float64):
    """Operator + FP64 FI Green with alpha = 0 attached (``adaptive_green_scaling/study.py:61-77``)."""
    op = Operator(rho, None, lam, mu, device, dtype).enable_green()
    set_alpha(op, 0.0)
    return op


def diagonal(op):
    """Nodal diag(K) (1, 3N) (``adaptive_green_scaling/study.py:80-88``)."""
    if op._diagonal is None:
        d = torch.empty((1, op.mesh.n_dofs), dtype=torch.float64, device=op.device)
        s = torch.empty_like(d)
        ws = torch.empty(_lib.lib().pgj_scale_ws_bytes(op.grid) // 8 + 1, dtype=torch.float64, device=op.device)
        _lib.call('pgj_diagonal_scale', op.grid, _lib.ptr(op._material_field), 0.0, _lib.ptr(d), _lib.ptr(s),
                  _lib.ptr(ws), _lib.stream())
        op._diagonal = d
    return op._diagonal


def scale(op, alpha):
    """s = exp(-alpha/2 (log d - mean log d)) (1, 3N) (``resident.py:170-174``)."""
    s = torch.empty((1, op.mesh.n_dofs), dtype=torch.float64, device=op.device)
    ws = torch.empty(_lib.lib().pgj_scale_ws_bytes(op.grid) // 8 + 1, dtype=torch.float64, device=op.device)
    _lib.call('pgj_diagonal_scale', op.grid, _lib.ptr(op._material_field), float(alpha), None, _lib.ptr(s),
              _lib.ptr(ws), _lib.stream())
    return s


def set_alpha(op, alpha):
    """Attach M_alpha^-1 r = P(s G(s P r)) as ``op.apply_precond`` (``adaptive_green_scaling/study.py:91-101``)."""
    if alpha == 0:
        op._scale = op._scale_hat = None
        op.apply_precond = lambda r: op.green_precond(r)
    else:
        if op._log_diagonal is None:
            q = diagonal(op).log()
            op._log_diagonal = q - q.mean()   # the attribute the reference attaches (study.py:96-98)
        s = scale(op, alpha)
        op._scale, op._scale_hat = s, op.scale_spectrum(s)
        op.apply_precond = lambda r: op.green_precond(r, op._scale, op._scale_hat)
    op._alpha = float(alpha)
    return op
