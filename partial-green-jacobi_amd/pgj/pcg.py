"""Operator-agnostic batched best-iterate PCG on libpgj's fused vector kernels.

``solve_pcg_batched`` keeps the signature, argument meaning, return value and
error behaviour of ``solver/src/runtime/fi/source/femlib/batched/pcg.py:20-252``:
``op`` provides ``apply`` ((B, n) -> (B, n)) and ``apply_precond`` (else
Jacobi from ``op.diag()``, ``pcg.py:287-295``); the result is the lowest
residual iterate per batch, the iteration count at the global stop and the
per-batch relative residuals. Every vector update and reduction runs in libpgj
(``pgj_pcg_*``): reductions use a fixed order, so results do not depend on the
launch. The update modes differ as in the reference:

* ``eager``, ``compiled`` with stride 1: best iterates keep improving until the
  host sees max(res) < tol at a check; the current residuals are returned.
* ``compiled`` with stride > 1: the device latches the first iteration where
  every residual is below tol; best iterates freeze there, and the latched
  iteration and best residuals are returned (``pcg.py:145-224``). With
  ``convergence_check_adaptive_factor`` the host switches to per-iteration
  checks once every residual is below factor * tol (``pcg.py:689-713``).
* ``triton``, ``cuda``: Jacobi only, z = M^-1 r fused into the r.z reduction
  (``pcg.py:297-455``).

Tensors must be FP64 (or FP32 throughout, for the inner solve of iterative
refinement: vectors FP32, sums and scalars FP64) and on the HIP device; there
is no host fallback.
"""
import math
import operator as _operator

import numpy as np
import torch

from . import _lib

_MODES = ('eager', 'compiled', 'triton', 'cuda')


def _validate_convergence_check_stride(value):
    try:
        stride = _operator.index(value)
    except TypeError as exc:
        raise ValueError('convergence_check_stride must be a positive integer') from exc
    if stride < 1:
        raise ValueError('convergence_check_stride must be at least one')
    return stride


def _validate_adaptive_factor(value):
    try:
        factor = float(value)
    except (TypeError, ValueError) as exc:
        raise ValueError('convergence_check_adaptive_factor must be finite and greater than one') from exc
    if not math.isfinite(factor) or factor <= 1.0:
        raise ValueError('convergence_check_adaptive_factor must be finite and greater than one')
    return factor


def _jacobi_m_inv(op):
    cache = getattr(op, '_M_inv_cache', None)
    if cache is None:
        cache = 1.0 / op.diag()
        op._M_inv_cache = cache
    return cache


def _make_precond(op):
    """r -> M^-1 r: ``op.apply_precond`` if present, else Jacobi (``pcg.py:287-295``)."""
    if hasattr(op, 'apply_precond'):
        return op.apply_precond
    cache = _jacobi_m_inv(op)
    return lambda r: cache * r


def _vector(t, shape, what, dtype=torch.float64):
    if not torch.is_tensor(t) or t.dtype != dtype or t.device.type != 'cuda':
        name = 'FP64' if dtype == torch.float64 else 'FP32'
        raise ValueError(f'{what} must be an {name} tensor on the HIP device')
    if tuple(t.shape) != tuple(shape):
        t = t.reshape(shape)
    return t.contiguous()


class _Workspace:
    """Device buffers of one solve: partial rows and per-batch scalars (FP64 for both vector precisions)."""

    def __init__(self, B, n, device, dtype=torch.float64):
        self.B, self.n = B, n
        self.pre = 'pgj_pcg_' if dtype == torch.float64 else 'pgj_pcg32_'
        self.ws = torch.empty(_lib.lib().pgj_pcg_ws_bytes(B, n) // 8 + 1, dtype=torch.float64, device=device)
        self.rhs_norm = torch.empty(B, dtype=torch.float64, device=device)
        self.rz = torch.empty(B, dtype=torch.float64, device=device)
        self.res = torch.full((B,), float('inf'), dtype=torch.float64, device=device)
        self.best_res = torch.full((B,), float('inf'), dtype=torch.float64, device=device)
        self.improved = torch.zeros(B, dtype=torch.uint8, device=device)
        self.state = torch.zeros(2, dtype=torch.int64, device=device)

    def start(self, rhs, ax, r):
        _lib.call(self.pre + 'start', self.B, self.n, _lib.ptr(rhs), _lib.ptr(ax), _lib.ptr(r), _lib.ptr(self.rhs_norm),
                  _lib.ptr(self.ws), _lib.stream())

    def init(self, r, z, p):
        _lib.call(self.pre + 'init', self.B, self.n, _lib.ptr(r), _lib.ptr(z), _lib.ptr(p), _lib.ptr(self.rz),
                  _lib.ptr(self.ws), _lib.stream())

    def pre_step(self, u, r, p, ap, best_u, tol, near_tol, iteration, latched):
        _lib.call(self.pre + 'pre_step', self.B, self.n, _lib.ptr(u), _lib.ptr(r), _lib.ptr(p), _lib.ptr(ap),
                  _lib.ptr(self.rz), _lib.ptr(self.rhs_norm), _lib.ptr(best_u), _lib.ptr(self.best_res),
                  _lib.ptr(self.res), _lib.ptr(self.improved), _lib.ptr(self.state), float(tol), float(near_tol),
                  int(iteration), int(latched), _lib.ptr(self.ws), _lib.stream())

    def post_step(self, r, z, p, minv=None):
        stride = 0 if minv is None or minv.numel() == self.n else self.n
        _lib.call(self.pre + 'post_step', self.B, self.n, _lib.ptr(r), _lib.ptr(z), _lib.ptr(minv), stride, _lib.ptr(p),
                  _lib.ptr(self.rz), _lib.ptr(self.ws), _lib.stream())

    def residual_norms(self, rhs, ax, out):
        """out = rhs - ax (ax None: rhs); returns |out_b| per batch (device tensor)."""
        B, n = self.B, self.n
        _lib.call(self.pre + 'start', B, n, _lib.ptr(rhs), _lib.ptr(ax), _lib.ptr(out), _lib.ptr(self.rhs_norm),
                  _lib.ptr(self.ws), _lib.stream())
        rr = torch.empty(B, dtype=torch.float64, device=rhs.device)
        copy = torch.empty_like(out)
        _lib.call(self.pre + 'init', B, n, _lib.ptr(out), _lib.ptr(out), _lib.ptr(copy), _lib.ptr(rr),
                  _lib.ptr(self.ws), _lib.stream())
        return rr.sqrt()


def _true_relative_residual(op, rhs, u, rhs_norm, work):
    """|rhs - A u| / rhs_norm per batch (``pcg.py:248-250``)."""
    ax = _vector(op.apply(u), rhs.shape, 'op.apply(u)', rhs.dtype)
    return work.residual_norms(rhs, ax, torch.empty_like(rhs)) / rhs_norm


def solve_pcg_batched(op, rhs, *, tol=1e-6, max_iter=2000, verbose=False, log_every=50, update_mode='eager',
                      x0=None, recycle_callback=None, recycle_stride=25, convergence_check_stride=1,
                      convergence_check_adaptive_factor=None, verify_final_residual=False):
    """Batched best-iterate PCG (``pcg.py:20-245``); returns (best_u (B, n), n_iter, rel_res (B,))."""
    convergence_check_stride = _validate_convergence_check_stride(convergence_check_stride)
    if update_mode not in _MODES:
        raise ValueError(f'unknown PCG update mode: {update_mode!r}')
    jacobi = update_mode in ('triton', 'cuda')
    if jacobi and getattr(op, '_fft_enabled', False):
        raise ValueError(f'update_mode={update_mode!r} currently supports Jacobi preconditioning only')
    if update_mode == 'cuda' and rhs.device.type != 'cuda':
        raise ValueError("update_mode='cuda' requires CUDA solver tensors")
    apply_m_inv = _make_precond(op)
    B, n = rhs.shape
    device = rhs.device
    adaptive = convergence_check_adaptive_factor is not None
    if adaptive:
        factor = _validate_adaptive_factor(convergence_check_adaptive_factor)
        if not (update_mode == 'compiled' and device.type == 'cuda' and convergence_check_stride > 1):
            raise ValueError("adaptive convergence checks require update_mode='compiled', "
                             'CUDA tensors, and convergence_check_stride > 1')
    dt = rhs.dtype if torch.is_tensor(rhs) and rhs.dtype == torch.float32 else torch.float64
    rhs = _vector(rhs, (B, n), 'rhs', dt)
    work = _Workspace(B, n, device, dt)

    if x0 is None:
        u = torch.zeros_like(rhs)
        r = torch.empty_like(rhs)
        work.start(rhs, None, r)
    else:
        u = _vector(x0, (B, n), 'x0', dt).clone()
        r = torch.empty_like(rhs)
        work.start(rhs, _vector(op.apply(u), (B, n), 'op.apply(x0)', dt), r)
    rhs_norm = work.rhs_norm.clone()

    z = _vector(apply_m_inv(r), (B, n), 'apply_precond(r)', dt)
    p = torch.empty_like(rhs)
    work.init(r, z, p)
    minv = None
    if jacobi:
        minv = _jacobi_m_inv(op)
        minv = _vector(minv, minv.shape, 'Jacobi diagonal inverse', dt)
        if minv.numel() not in (n, B * n):
            minv = torch.broadcast_to(minv, (B, n)).contiguous()
        z = torch.empty_like(rhs)
    best_u = u.clone()
    latched = update_mode == 'compiled' and convergence_check_stride > 1
    near_tol = factor * tol if adaptive else 0.0
    host_check_stride = convergence_check_stride
    recycle = recycle_callback if update_mode != 'triton' else None
    recycle_stride = max(int(recycle_stride), 1)
    tag = {'eager': '', 'compiled': '', 'triton': ':Triton', 'cuda': ':CUDA'}[update_mode]

    n_iter_final = max_iter
    for it in range(max_iter):
        ap = _vector(op.apply(p), (B, n), 'op.apply(p)', dt)
        if recycle is not None and it % recycle_stride == 0:
            recycle(it, p.clone(), ap)
        work.pre_step(u, r, p, ap, best_u, tol, near_tol, it + 1, latched)

        if verbose and ((it + 1) % log_every == 0 or it == 0):
            print(f'    [PCG-batched{tag}] iter {it + 1:5d}  max_res={work.res.max().item():.3e}  '
                  f'min_res={work.res.min().item():.3e}')

        check_now = it == 0 or (it + 1) % host_check_stride == 0 or (it + 1) == max_iter
        logical = 0
        if check_now:
            if latched:
                stop, near = (int(v) for v in work.state.tolist())
                if adaptive and stop == 0 and near:
                    host_check_stride = 1
                logical = stop
            elif work.res.max().item() < tol:
                logical = it + 1
        if logical > 0:
            if verbose:
                print(f'    [PCG-batched{tag}] iter {logical:5d}  all batches converged')
            if verify_final_residual:
                out = _true_relative_residual(op, rhs, best_u, rhs_norm, work)
            else:
                out = (work.best_res if latched else work.res).clone()
            return best_u, logical, out
        if (it + 1) == max_iter:
            break
        if jacobi:
            work.post_step(r, z, p, minv)
        else:
            z = _vector(apply_m_inv(r), (B, n), 'apply_precond(r)', dt)
            work.post_step(r, z, p)

    if verbose:
        print(f'    [PCG-batched{tag}] max_iter={max_iter} reached, max_res={work.res.max().item():.3e}')
    if verify_final_residual:
        return best_u, n_iter_final, _true_relative_residual(op, rhs, best_u, rhs_norm, work)
    return best_u, n_iter_final, work.best_res.clone()


def warmup_compiled_pcg_updates(op, rhs, *, tol=1e-6, convergence_check_stride=1,
                                convergence_check_adaptive_factor=None):
    """Argument checks of ``pcg.py:539-618``; the fused kernels are prebuilt, so nothing is compiled."""
    stride = _validate_convergence_check_stride(convergence_check_stride)
    _make_precond(op)
    if convergence_check_adaptive_factor is not None:
        _validate_adaptive_factor(convergence_check_adaptive_factor)
        if rhs.device.type != 'cuda' or stride <= 1:
            raise ValueError('adaptive convergence checks require CUDA tensors and convergence_check_stride > 1')
    _lib.lib()


def certified_solve(op, rhs, x0=None, rtol=1e-8, maxiter=4000):
    """PCG with up to two restarts until |rhs - A u| <= max(1e-13/sqrt(nx), rtol |rhs|).

    Mirrors ``Backend.solve`` (``exp/asg_topopt_20261103/backend.py:122-141``):
    compiled updates with host checks every 4 iterations, zero-mean projection
    of every restart point, the true residual measured after each attempt.
    Returns (u, info).
    """
    rhs = _vector(rhs, rhs.shape, 'rhs')
    B, n = rhs.shape
    work = _Workspace(B, n, rhs.device)
    scratch = torch.empty_like(rhs)

    def true_absolute(u):
        ax = _vector(op.apply(u), (B, n), 'op.apply(u)')
        norms = work.residual_norms(rhs, ax, scratch)
        return math.sqrt(float((norms * norms).sum()))

    rhs_norms = work.residual_norms(rhs, None, scratch)
    norm = math.sqrt(float((rhs_norms * rhs_norms).sum()))
    atol = 1e-13 / math.sqrt(op.mesh.Nx)
    target = max(atol, rtol * norm)
    u = torch.zeros_like(rhs) if x0 is None else _vector(op.project_zero_mean(x0), (B, n), 'x0')
    absolute = true_absolute(u)
    iterations = restarts = 0
    if absolute > target:
        while iterations < maxiter:
            u, it, _ = solve_pcg_batched(op, rhs, x0=u, tol=target / max(norm, 1e-30), max_iter=maxiter - iterations,
                                         update_mode='compiled', convergence_check_stride=4)
            iterations += it
            u = op.project_zero_mean(u)
            absolute = true_absolute(u)
            if absolute <= target or not np.isfinite(absolute) or restarts >= 2:
                break
            restarts += 1
    info = dict(iterations=iterations, restarts=restarts, rhs_norm=norm, absolute_residual=absolute,
                relative_residual=absolute / max(norm, 1e-300), threshold=target,
                converged=bool(np.isfinite(absolute) and absolute <= target))
    return u, info
