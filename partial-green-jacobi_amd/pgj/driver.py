"""Average-stress effective tensor over the six unit strains: the second C^H route.

``compute_effective_stiffness_batched`` keeps the signature, defaults, return
value and info keys of ``solver/src/runtime/fi/source/femlib/batched/driver.py:13-75``:
for each unit Voigt strain E_J it solves K u = f(E_J) with
``solve_pcg_batched`` (the drop-in PCG on libpgj), projects u to zero mean and
takes the mean stress <a C0 (E_J + B u)> (``pgj_hex8_avg_stress``) as column J;
the result is symmetrised. It cross-checks the energy C^H of the resident
solver (``effective_tensor``), which polarises element energies instead.
"""
import os
import time

import numpy as np

from .pcg import solve_pcg_batched


def compute_effective_stiffness_batched(op, *, tol=1e-6, max_iter=2000, verbose=False, pcg_update_mode='eager',
                                        convergence_check_stride=1, convergence_check_adaptive_factor=None,
                                        verify_final_residual=False):
    """Six load cases over the ``op.B`` samples (``driver.py:13-75``).

    Returns ``(C (B, 6, 6) numpy, info)`` with ``info[b][J]`` holding ``iters``, ``rel_res``, ``converged``,
    ``residual_verified`` and ``elapsed_s`` (the load case's wall time averaged over the samples).
    """
    B = op.B
    C = np.zeros((B, 6, 6), dtype=np.float64)
    info = [[None] * 6 for _ in range(B)]
    log_every = int(os.environ.get('MFREE_PCG_LOG_EVERY', '50'))
    for J in range(6):
        E = np.zeros(6, dtype=np.float64)
        E[J] = 1.0
        t0 = time.time()
        rhs = op.compute_rhs(E)
        u, n_iter, rel_res = solve_pcg_batched(
            op, rhs, tol=tol, max_iter=max_iter, verbose=verbose, log_every=log_every, update_mode=pcg_update_mode,
            convergence_check_stride=convergence_check_stride,
            convergence_check_adaptive_factor=convergence_check_adaptive_factor,
            verify_final_residual=verify_final_residual)
        u = op.project_zero_mean(u)
        C[:, :, J] = op.compute_avg_stress(u, E)
        elapsed = time.time() - t0
        res = rel_res.detach().cpu().numpy().astype(np.float64)
        for b in range(B):
            info[b][J] = {'iters': int(n_iter), 'rel_res': float(res[b]), 'converged': bool(res[b] < tol),
                          'residual_verified': bool(verify_final_residual), 'elapsed_s': float(elapsed / B)}
        if verbose:
            print(f'  [Batched] load case {J + 1}/6: iters={n_iter}, max_res={res.max():.3e}, {elapsed:.2f}s')
    C = 0.5 * (C + np.transpose(C, (0, 2, 1)))
    return C, info
