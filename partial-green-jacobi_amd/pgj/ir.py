"""Batched mixed-precision iterative refinement over the generic PCG.

Keeps the signatures, defaults, return values, info keys and errors of
``solver/src/runtime/fi/source/femlib/batched/ir.py``: the outer loop measures
r = f - K u with ``op_outer`` (FP64), solves K du = r loosely with
``solve_pcg_batched`` on ``op_inner`` (FP32, e.g. ``pgj.Operator(...,
dtype=torch.float32)``), adds du in FP64 and projects to zero mean; every batch
keeps its best iterate by its own outer residual. The adaptive variant tightens
the inner tolerance when the outer residual stagnates and stops early when it
stalls at the tightest one. The outer residual and its norms run on libpgj's
FP64 vector kernels; the inner solve on its FP32 ones.
"""
import time

import numpy as np
import torch

from .pcg import _Workspace, _vector, solve_pcg_batched


class _Outer:
    """|rhs - K u| / |rhs| per batch on the FP64 vector kernels."""

    def __init__(self, op_outer, rhs):
        self.op = op_outer
        self.rhs = _vector(rhs, rhs.shape, 'rhs')
        B, n = self.rhs.shape
        self.work = _Workspace(B, n, rhs.device)
        self.scratch = torch.empty_like(self.rhs)
        self.rhs_norm = self.work.residual_norms(self.rhs, None, self.scratch).clamp(min=1e-30)

    def residual(self, u):
        ax = _vector(self.op.apply(u), self.rhs.shape, 'op_outer.apply(u)')
        rel = self.work.residual_norms(self.rhs, ax, self.scratch) / self.rhs_norm
        return self.scratch, rel


def _check_batches(op_outer, op_inner):
    if op_outer.B != op_inner.B:
        raise AssertionError(f'batch size mismatch: outer={op_outer.B}, inner={op_inner.B}')


def solve_ir_batched(op_outer, op_inner, rhs, *, outer_tol=1e-8, inner_tol=0.1, max_outer=30, max_inner=80,
                     inner_pcg_update_mode='eager', inner_convergence_check_stride=1,
                     inner_verify_final_residual=False, verbose=False):
    """Fixed inner tolerance (``ir.py:23-107``); returns (best_u (B, n), info)."""
    _check_batches(op_outer, op_inner)
    B = op_outer.B
    in_dtype = op_inner.dtype
    outer = _Outer(op_outer, rhs)
    u = torch.zeros_like(outer.rhs)
    best_res = torch.full((B,), float('inf'), device=u.device, dtype=torch.float64)
    best_u = u.clone()
    total_inner, last_k = 0, 0
    for k in range(max_outer):
        r_out, rel_res = outer.residual(u)
        last_k = k
        improved = rel_res < best_res
        if bool(improved.any()):
            best_u = torch.where(improved.unsqueeze(-1), u, best_u)
            best_res = torch.where(improved, rel_res, best_res)
        if verbose:
            print(f'    [IR-batched] outer {k:3d}  max_res={rel_res.max().item():.3e}  '
                  f'min_res={rel_res.min().item():.3e}')
        if rel_res.max().item() < outer_tol:
            if verbose:
                print(f'    [IR-batched] outer {k:3d}  all batches converged')
            break
        du_in, n_in, _ = solve_pcg_batched(op_inner, r_out.to(in_dtype), tol=inner_tol, max_iter=max_inner,
                                           verbose=False, update_mode=inner_pcg_update_mode,
                                           convergence_check_stride=inner_convergence_check_stride,
                                           verify_final_residual=inner_verify_final_residual)
        total_inner += int(n_in)
        u = op_outer.project_zero_mean(u + du_in.to(torch.float64))
    converged = bool(best_res.max().item() < outer_tol)
    info = {'outer_iters': int(last_k + 1), 'inner_iters': int(total_inner),
            'rel_res': best_res.detach().cpu().numpy().astype(np.float64), 'converged': converged,
            'stop_reason': 'converged' if converged else 'max_outer',
            'inner_pcg_update_mode': str(inner_pcg_update_mode)}
    return best_u, info


def solve_ir_adaptive_batched(op_outer, op_inner, rhs, *, outer_tol=1e-8, inner_tols=(0.1, 0.03, 0.01),
                              max_outer=30, max_inner=80, adaptive_max_inner=160, stagnation_factor=0.3,
                              stall_factor=0.9, stall_patience=3, inner_pcg_update_mode='eager',
                              inner_convergence_check_stride=1, inner_verify_final_residual=False, verbose=False):
    """Adaptive inner tolerance (``ir.py:110-247``); returns (best_u (B, n), info) with a per-step history."""
    _check_batches(op_outer, op_inner)
    inner_tols = tuple(float(t) for t in inner_tols)
    if not inner_tols:
        raise ValueError('inner_tols must contain at least one tolerance')
    if any(t <= 0.0 for t in inner_tols):
        raise ValueError(f'inner_tols must be positive, got {inner_tols}')
    if max_outer < 1:
        raise ValueError('max_outer must be >= 1')
    if max_inner < 1 or adaptive_max_inner < 1:
        raise ValueError('max_inner and adaptive_max_inner must be >= 1')
    B = op_outer.B
    in_dtype = op_inner.dtype
    outer = _Outer(op_outer, rhs)
    u = torch.zeros_like(outer.rhs)
    best_res = torch.full((B,), float('inf'), device=u.device, dtype=torch.float64)
    best_u = u.clone()
    total_inner, last_k = 0, 0
    prev_outer_max = None
    tol_index = 0
    current_max_inner = int(max_inner)
    stall_count = 0
    stop_reason = 'max_outer'
    history = []
    for k in range(max_outer):
        r_out, rel_res = outer.residual(u)
        outer_max, outer_min = float(rel_res.max().item()), float(rel_res.min().item())
        last_k = k
        improved = rel_res < best_res
        if bool(improved.any()):
            best_u = torch.where(improved.unsqueeze(-1), u, best_u)
            best_res = torch.where(improved, rel_res, best_res)
        if verbose:
            print(f'    [IR-adaptive] outer {k:3d}  max_res={outer_max:.3e}  min_res={outer_min:.3e}  '
                  f'inner_tol={inner_tols[tol_index]:.2e}')
        if outer_max < outer_tol:
            stop_reason = 'converged'
            break
        tightened = False
        if prev_outer_max is not None and outer_max > float(stagnation_factor) * prev_outer_max:
            if tol_index < len(inner_tols) - 1:
                tol_index += 1
                current_max_inner = max(current_max_inner, int(adaptive_max_inner))
                tightened = True
                stall_count = 0
            elif outer_max > float(stall_factor) * prev_outer_max:
                stall_count += 1
            else:
                stall_count = 0
        else:
            stall_count = 0
        if stall_count >= int(stall_patience):
            stop_reason = 'stalled'
            if verbose:
                print(f'    [IR-adaptive] outer {k:3d}  stalled; caller should fall back to fp64')
            break
        inner_tol = inner_tols[tol_index]
        du_in, n_in, inner_rel_res = solve_pcg_batched(
            op_inner, r_out.to(in_dtype), tol=inner_tol, max_iter=current_max_inner, verbose=False,
            update_mode=inner_pcg_update_mode, convergence_check_stride=inner_convergence_check_stride,
            verify_final_residual=inner_verify_final_residual)
        total_inner += int(n_in)
        history.append({'outer_iter': int(k), 'rel_res_max': outer_max, 'rel_res_min': outer_min,
                        'inner_tol': float(inner_tol), 'inner_max_iter': int(current_max_inner),
                        'inner_iters': int(n_in), 'inner_rel_res_max': float(inner_rel_res.max().item()),
                        'tightened': bool(tightened), 'stall_count': int(stall_count)})
        u = op_outer.project_zero_mean(u + du_in.to(torch.float64))
        prev_outer_max = outer_max
    converged = bool(best_res.max().item() < outer_tol)
    if converged:
        stop_reason = 'converged'
    info = {'outer_iters': int(last_k + 1), 'inner_iters': int(total_inner),
            'rel_res': best_res.detach().cpu().numpy().astype(np.float64), 'converged': converged,
            'stop_reason': stop_reason, 'adaptive': True, 'inner_tols': [float(v) for v in inner_tols],
            'final_inner_tol': float(inner_tols[tol_index]), 'max_inner_used': int(current_max_inner),
            'inner_pcg_update_mode': str(inner_pcg_update_mode), 'history': history}
    return best_u, info


def _backend_keys(op_outer, op_inner):
    return {'outer_apply_backend': getattr(op_outer, 'apply_backend', 'unspecified'),
            'inner_apply_backend': getattr(op_inner, 'apply_backend', 'unspecified'),
            'outer_cuda_configuration': getattr(op_outer, 'cuda_apply_configuration', None),
            'inner_cuda_configuration': getattr(op_inner, 'cuda_apply_configuration', None)}


def _ir_driver(op_outer, op_inner, solve, extra, verbose):
    B = op_outer.B
    C = np.zeros((B, 6, 6), dtype=np.float64)
    info = [[None] * 6 for _ in range(B)]
    for J in range(6):
        E = np.zeros(6, dtype=np.float64)
        E[J] = 1.0
        t0 = time.time()
        rhs = op_outer.compute_rhs(E)
        u, step = solve(rhs)
        u = op_outer.project_zero_mean(u)
        C[:, :, J] = op_outer.compute_avg_stress(u, E)
        elapsed = time.time() - t0
        for b in range(B):
            entry = {'outer_iters': step['outer_iters'], 'inner_iters': step['inner_iters'],
                     'rel_res': float(step['rel_res'][b]), 'elapsed_s': float(elapsed / B)}
            entry.update({k: step[k] for k in extra})
            entry.update(_backend_keys(op_outer, op_inner))
            if 'history' in step:
                entry['history'] = step['history']
            info[b][J] = entry
        if verbose:
            print(f"  [IR-batched] load case {J + 1}/6: outer={step['outer_iters']}, inner={step['inner_iters']}, "
                  f"max_res={step['rel_res'].max():.3e}, {elapsed:.2f}s")
    return 0.5 * (C + np.transpose(C, (0, 2, 1))), info


def compute_effective_stiffness_ir_batched(op_outer, op_inner, *, outer_tol=1e-8, inner_tol=0.1, max_outer=30,
                                           max_inner=80, inner_pcg_update_mode='eager',
                                           inner_convergence_check_stride=1, inner_verify_final_residual=False,
                                           verbose=False):
    """Six unit strains through ``solve_ir_batched`` and the mean stress (``ir.py:250-318``)."""
    def solve(rhs):
        return solve_ir_batched(op_outer, op_inner, rhs, outer_tol=outer_tol, inner_tol=inner_tol,
                                max_outer=max_outer, max_inner=max_inner,
                                inner_pcg_update_mode=inner_pcg_update_mode,
                                inner_convergence_check_stride=inner_convergence_check_stride,
                                inner_verify_final_residual=inner_verify_final_residual, verbose=verbose)
    return _ir_driver(op_outer, op_inner, solve, ('inner_pcg_update_mode',), verbose)


def compute_effective_stiffness_ir_adaptive_batched(op_outer, op_inner, *, outer_tol=1e-8,
                                                    inner_tols=(0.1, 0.03, 0.01), max_outer=30, max_inner=80,
                                                    adaptive_max_inner=160, stagnation_factor=0.3,
                                                    stall_factor=0.9, stall_patience=3,
                                                    inner_pcg_update_mode='eager', inner_convergence_check_stride=1,
                                                    inner_verify_final_residual=False, verbose=False):
    """Six unit strains through ``solve_ir_adaptive_batched`` (``ir.py:321-400``)."""
    def solve(rhs):
        return solve_ir_adaptive_batched(
            op_outer, op_inner, rhs, outer_tol=outer_tol, inner_tols=inner_tols, max_outer=max_outer,
            max_inner=max_inner, adaptive_max_inner=adaptive_max_inner, stagnation_factor=stagnation_factor,
            stall_factor=stall_factor, stall_patience=stall_patience, inner_pcg_update_mode=inner_pcg_update_mode,
            inner_convergence_check_stride=inner_convergence_check_stride,
            inner_verify_final_residual=inner_verify_final_residual, verbose=verbose)
    extra = ('converged', 'stop_reason', 'adaptive', 'inner_tols', 'final_inner_tol', 'max_inner_used',
             'inner_pcg_update_mode')
    return _ir_driver(op_outer, op_inner, solve, extra, verbose)
