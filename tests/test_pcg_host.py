"""Host-side behaviour of the solve_pcg_batched drop-in (CPU, no device work).

Argument validation mirrors solver/src/runtime/fi/source/femlib/batched/pcg.py:44-75,
:260-284; the oracle's adaptive host checks are pinned against the fixed-stride path.
"""
import numpy as np
import pytest
import torch

from oracle import pcg as opcg
from pgj import pcg


class _Jacobi:
    def __init__(self, diag):
        self._d = diag

    def apply(self, u):
        return u * self._d

    def diag(self):
        return self._d


class _FFT(_Jacobi):
    _fft_enabled = True

    def apply_precond(self, r):
        return r


@pytest.mark.parametrize('stride', [0, -3, 1.5, '4', None])
def test_stride_validation(stride):
    with pytest.raises(ValueError, match='convergence_check_stride'):
        pcg.solve_pcg_batched(_Jacobi(torch.ones(4)), torch.ones(1, 4), convergence_check_stride=stride)


def test_unknown_mode():
    with pytest.raises(ValueError, match='unknown PCG update mode'):
        pcg.solve_pcg_batched(_Jacobi(torch.ones(4)), torch.ones(1, 4), update_mode='fused')


@pytest.mark.parametrize('mode', ['triton', 'cuda'])
def test_jacobi_modes_reject_spectral_preconditioner(mode):
    with pytest.raises(ValueError, match='Jacobi preconditioning only'):
        pcg.solve_pcg_batched(_FFT(torch.ones(4)), torch.ones(1, 4), update_mode=mode)


def test_cuda_mode_requires_device_tensors():
    with pytest.raises(ValueError, match='requires CUDA'):
        pcg.solve_pcg_batched(_Jacobi(torch.ones(4)), torch.ones(1, 4), update_mode='cuda')


@pytest.mark.parametrize('factor', [1.0, 0.5, float('nan'), float('inf'), 'x'])
def test_adaptive_factor_validation(factor):
    with pytest.raises(ValueError, match='adaptive_factor'):
        pcg.solve_pcg_batched(_Jacobi(torch.ones(4)), torch.ones(1, 4), update_mode='compiled',
                              convergence_check_stride=4, convergence_check_adaptive_factor=factor)


def test_adaptive_requires_latched_device_path():
    with pytest.raises(ValueError, match='adaptive convergence checks require'):
        pcg.solve_pcg_batched(_Jacobi(torch.ones(4)), torch.ones(1, 4), update_mode='compiled',
                              convergence_check_stride=1, convergence_check_adaptive_factor=10.0)


def test_host_tensors_are_refused():
    """No host fallback: CPU tensors fail loudly after validation."""
    with pytest.raises(ValueError, match='HIP device'):
        pcg.solve_pcg_batched(_Jacobi(torch.ones(4, dtype=torch.float64)), torch.ones(1, 4, dtype=torch.float64))


def _spd(n, seed, cond=200.0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (q * np.geomspace(1.0, cond, n)) @ q.T


def test_oracle_adaptive_stops_at_latched_iteration():
    """Adaptive checks return the same latched iteration and iterate as stride 1, never later than stride 8."""
    A = _spd(60, 3)
    rhs = np.random.default_rng(4).standard_normal((3, 60))
    d = np.diag(A)
    apply = lambda u: u @ A.T
    pre = lambda r: r / d
    ref = opcg.solve_pcg_batched(apply, pre, rhs, tol=1e-9, max_iter=500, update_mode='compiled',
                                 convergence_check_stride=2)
    ad = opcg.solve_pcg_batched(apply, pre, rhs, tol=1e-9, max_iter=500, update_mode='compiled',
                                convergence_check_stride=8, adaptive_factor=1e3)
    fixed = opcg.solve_pcg_batched(apply, pre, rhs, tol=1e-9, max_iter=500, update_mode='compiled',
                                   convergence_check_stride=8)
    assert ad[1] == ref[1] == fixed[1]          # the latched iteration itself does not depend on checks
    np.testing.assert_array_equal(ad[0], ref[0])
    np.testing.assert_array_equal(ad[2], fixed[2])
    # the true relative residual of the returned iterate meets tol
    rr = np.linalg.norm(rhs - apply(ad[0]), axis=1) / np.linalg.norm(rhs, axis=1)
    assert rr.max() < 1e-9 * 1.01


def test_oracle_adaptive_requires_latched():
    with pytest.raises(ValueError):
        opcg.solve_pcg_batched(lambda u: u, lambda r: r, np.ones((1, 3)), update_mode='eager',
                               adaptive_factor=10.0)


class _Shell:
    def __init__(self, B, dtype=torch.float64):
        self.B, self.dtype = B, dtype


def test_ir_argument_errors_before_device_work():
    from pgj import ir
    rhs = torch.zeros(1, 6, dtype=torch.float64)
    with pytest.raises(AssertionError, match='batch size mismatch'):
        ir.solve_ir_batched(_Shell(1), _Shell(2, torch.float32), rhs)
    with pytest.raises(ValueError, match='at least one tolerance'):
        ir.solve_ir_adaptive_batched(_Shell(1), _Shell(1), rhs, inner_tols=())
    with pytest.raises(ValueError, match='must be positive'):
        ir.solve_ir_adaptive_batched(_Shell(1), _Shell(1), rhs, inner_tols=(0.1, 0.0))
    with pytest.raises(ValueError, match='max_outer'):
        ir.solve_ir_adaptive_batched(_Shell(1), _Shell(1), rhs, max_outer=0)
    with pytest.raises(ValueError, match='adaptive_max_inner'):
        ir.solve_ir_adaptive_batched(_Shell(1), _Shell(1), rhs, adaptive_max_inner=0)
