"""Multi-load PCG recurrences and certification, restated in numpy (test infrastructure only).

``ResidentPCG`` follows ``exp/asg_gpu_resident_20261103/resident.py:39-58``
(update_ur/update_p), ``:119-136`` (initialize/iteration/certify) and
``:176-209`` (load modes, targets, restarts, RuntimeError), with the 6-load
grouping of ``exp/asg_sctop_reproduction_20261105/reproduction.py:144-151``.
``solve_pcg_batched`` follows ``solver/src/runtime/fi/source/femlib/batched/pcg.py:20-252``
(eager and compiled/latched update modes, best-iterate return) and
``backend_solve`` follows ``exp/asg_topopt_20261103/backend.py:123-144`` and
``solve_ir_batched`` follows ``solver/src/runtime/fi/source/femlib/batched/ir.py:23-247``.
Pinned in ``tests/test_oracle_pcg.py`` against dense direct solves, exact
zero-RHS behaviour, the iteration cap and the restart/RuntimeError paths; the
refinement against the dense solution and the closed-form contraction of an
inner operator c K (outer residual ratio 1 - 1/c: the tightening and stall
schedule follows from it).
"""
import math

import numpy as np

from . import green, hex8

CAP = 4000


class ResidentPCG:
    """CPU restatement of ``Resident``/``PaperResident`` solve semantics.

    ``a`` is the SIMP field (n,n,n); ``s`` the (1,3N) scale (None: Green);
    ``loads`` the (L,6) macro strains; ``batch`` 1 (serial) or the group size.
    """

    def __init__(self, a, C, loads, batch=1, s=None, lam_mu=None, chunk=4):
        self.a = np.asarray(a, float)
        self.grid = self.a.shape
        self.n = self.grid[2]
        self.voxel = hex8.voxel_of(self.grid)
        self.C = np.asarray(C, float)
        lam, mu = (self.C[0, 1], self.C[3, 3]) if lam_mu is None else lam_mu
        self.G6 = green.symbol(self.grid, self.voxel, lam, mu)
        self.loads = np.atleast_2d(np.asarray(loads, float))
        self.batch = batch
        self.s = s
        self.chunk = chunk
        N = int(np.prod(self.grid))
        self.solutions = np.zeros((len(self.loads), 3 * N))

    def K(self, p, active):
        return hex8.apply(self.a, p, self.C, self.voxel, active)

    def true_K(self, u):
        return hex8.apply(self.a, u, self.C, self.voxel)

    def B(self, r):
        return green.precond(r, self.G6, self.grid, self.s)

    def rhs(self, macro):
        return hex8.rhs(self.a, macro, self.C, self.voxel)

    def solve(self, rtol=1e-8):
        infos, checks = [], 0
        for start in range(0, len(self.loads), self.batch):
            idx = list(range(start, min(start + self.batch, len(self.loads))))
            u, info, count = self.solve_batch(idx, self.solutions[idx].copy(), rtol)
            self.solutions[idx] = u
            infos.extend(info)
            checks += count
        return infos, checks

    def solve_batch(self, load_indices, u, rtol):
        L = len(load_indices)
        iterations = np.zeros(L, np.int64)
        bad = np.zeros(L, bool)
        f = np.concatenate([self.rhs(self.loads[j]) for j in load_indices])
        norm = np.sqrt((f * f).sum(1))
        target2 = np.maximum(norm * rtol, np.full_like(norm, 1e-13 / math.sqrt(self.n))) ** 2
        restarts = checks = 0

        def certify(u):
            u = hex8.project(u)
            r = f - self.true_K(u)
            return u, r, np.sqrt((r * r).sum(1))

        u, r, absolute = certify(u)
        while True:
            active = ((r * r).sum(1) > target2) & (iterations < CAP) & ~bad
            z = self.B(r)
            p = np.where(active[:, None], z, 0.0)
            rz = (r * z).sum(1)
            while True:
                checks += 1
                if not active.any():
                    break
                for _ in range(self.chunk):
                    u, r, p, rz, active, iterations, bad = self.iteration(
                        u, r, p, rz, active, iterations, bad, target2)
            u, r, absolute = certify(u)
            ok = np.isfinite(absolute) & (absolute <= np.sqrt(target2)) & ~bad
            if ok.all() or restarts >= 2 or bad.any() or np.any(~ok & (iterations >= CAP)):
                break
            restarts += 1
        info = [dict(rhs_norm=float(norm[j]), absolute_residual=float(absolute[j]),
                     threshold=float(np.sqrt(target2[j])), iterations=int(iterations[j]),
                     relative_residual=float(absolute[j] / max(norm[j], 1e-300)),
                     converged=bool(ok[j]), restarts=restarts) for j in range(L)]
        if not ok.all():
            raise RuntimeError(('true residual certification failed', info))
        return u, info, checks

    def iteration(self, u, r, p, rz, active, iterations, bad, target2):
        q = self.K(p, active)
        with np.errstate(invalid='ignore', divide='ignore', over='ignore'):
            den = (p * q).sum(1)
            valid = np.isfinite(den) & (den > 0) & np.isfinite(rz) & (rz > 0)
            failed = active & ~valid
            alpha = np.where(active & valid, rz / np.where(valid, den, 1.0), 0.0)
            u = u + alpha[:, None] * p
            r = r - alpha[:, None] * q
            res2 = (r * r).sum(1)
            finite = np.isfinite(res2)
            it = iterations + active.astype(np.int64)
            keep = active & valid & finite & (res2 > target2) & (it < CAP)
            bad = bad | failed | (active & ~finite)
            active = keep
            z = self.B(r)
            newrz = (r * z).sum(1)
            beta = np.where(active, newrz / np.where(rz != 0, rz, 1.0), 0.0)
            p = np.where(active[:, None], z + beta[:, None] * p, 0.0)
        return u, r, p, newrz, active, it, bad


class HybridResidentPCG(ResidentPCG):
    """``ResidentPCG`` over a general-C21 material (``oracle.hybrid``): K and the rhs from the pool rows
    (``index`` None: row e for element e), the Green preconditioner of the isotropic medium ``lam_mu`` (the
    operator's reference medium) without scaling. The recurrence, certification and restarts are the
    resident solver's (``resident.py:39-58``, ``:186-209``)."""

    def __init__(self, pool, index, grid, voxel, loads, batch=1, lam_mu=None, chunk=4):
        from . import hybrid
        self._hybrid = hybrid
        self.pool = np.asarray(pool, float).reshape(-1, 21)
        self.index = None if index is None else np.asarray(index).reshape(-1)
        self.grid = tuple(int(v) for v in grid)
        self.n = self.grid[2]
        self.voxel = tuple(float(v) for v in voxel)
        self.G6 = green.symbol(self.grid, self.voxel, *lam_mu)
        self.loads = np.atleast_2d(np.asarray(loads, float))
        self.batch = batch
        self.s = None
        self.chunk = chunk
        self.solutions = np.zeros((len(self.loads), 3 * int(np.prod(self.grid))))

    def K(self, p, active):
        return self._hybrid.apply(self.pool, self.index, p, self.grid, self.voxel, active)

    def true_K(self, u):
        return self._hybrid.apply(self.pool, self.index, u, self.grid, self.voxel)

    def rhs(self, macro):
        return self._hybrid.rhs(self.pool, self.index, macro, self.grid, self.voxel)


def solve_pcg_batched(apply, apply_precond, rhs, *, tol=1e-6, max_iter=2000, x0=None,
                      update_mode='eager', convergence_check_stride=1, adaptive_factor=None):
    """Best-iterate batched PCG (``pcg.py:20-252``); returns (best_u, n_iter, rel_res).

    ``update_mode='compiled'`` with stride > 1 follows the device-latched path
    (``pcg.py:145-224``, ``:676-738``): best iterates stop updating once every
    batch is below ``tol``; the host checks at iteration 1, every ``stride``
    iterations and at ``max_iter``. ``adaptive_factor`` (latched path only)
    switches the host to per-iteration checks once every batch is below
    ``adaptive_factor * tol`` (``pcg.py:689-713``, ``:205-217``).
    """
    if update_mode not in ('eager', 'compiled'):
        raise ValueError(f'unknown PCG update mode: {update_mode!r}')
    if int(convergence_check_stride) != convergence_check_stride or convergence_check_stride < 1:
        raise ValueError('convergence_check_stride must be at least one')
    rhs = np.atleast_2d(np.asarray(rhs, float))
    B = rhs.shape[0]
    eps = 1e-30
    if x0 is None:
        u = np.zeros_like(rhs)
        r = rhs.copy()
    else:
        u = np.array(x0, float)
        r = rhs - apply(u)
    rhs_norm = np.maximum(np.sqrt((rhs * rhs).sum(1)), 1e-30)
    z = apply_precond(r)
    p = z.copy()
    rz = (r * z).sum(1)
    latched = update_mode == 'compiled' and convergence_check_stride > 1
    if adaptive_factor is not None and not latched:
        raise ValueError('adaptive convergence checks require the latched path')
    host_stride = convergence_check_stride
    near_latched = False
    best_res = np.full(B, np.inf)
    best_u = u.copy()
    stop_iter = 0
    n_iter_final = max_iter
    res = np.full(B, np.inf)
    for it in range(max_iter):
        Ap = apply(p)
        with np.errstate(invalid='ignore', divide='ignore', over='ignore'):
            pAp = (p * Ap).sum(1)
            bad = ~np.isfinite(pAp) | (pAp <= 0)
            alpha = rz / np.maximum(pAp, eps)
            alpha = np.where(bad, 0.0, alpha)
            u = u + alpha[:, None] * p
            r = r - alpha[:, None] * Ap
            res = np.sqrt((r * r).sum(1)) / rhs_norm
        improved = res < best_res
        if latched:
            improved &= stop_iter == 0
        best_u = np.where(improved[:, None], u, best_u)
        best_res = np.where(improved, res, best_res)
        if latched and stop_iter == 0 and np.all(res < tol):
            stop_iter = it + 1
        if adaptive_factor is not None and np.all(res < adaptive_factor * tol):
            near_latched = True
        check_now = it == 0 or (it + 1) % host_stride == 0 or (it + 1) == max_iter
        logical = 0
        if check_now:
            if latched:
                if adaptive_factor is not None and stop_iter == 0 and near_latched:
                    host_stride = 1
                logical = stop_iter
            elif res.max() < tol:
                logical = it + 1
        if logical > 0:
            return best_u, logical, (best_res if latched else res)
        if (it + 1) == max_iter:
            break
        z = apply_precond(r)
        rz_new = (r * z).sum(1)
        beta = rz_new / np.maximum(rz, eps)
        p = z + beta[:, None] * p
        rz = rz_new
    return best_u, n_iter_final, best_res


def solve_ir_batched(apply_outer, apply_inner, precond_inner, project, rhs, *, outer_tol=1e-8, inner_tol=0.1,
                     max_outer=30, max_inner=80, inner_tols=None, adaptive_max_inner=160, stagnation_factor=0.3,
                     stall_factor=0.9, stall_patience=3):
    """Iterative refinement (``solver/src/runtime/fi/source/femlib/batched/ir.py:23-107``; with
    ``inner_tols`` the adaptive variant ``ir.py:110-247``). Inner solves are FP64 eager
    ``solve_pcg_batched`` on ``apply_inner``. Returns (best_u, info) with ``outer_iters``,
    ``inner_iters``, ``rel_res``, ``stop_reason`` and (adaptive) the per-step ``history``."""
    rhs = np.atleast_2d(np.asarray(rhs, float))
    B = rhs.shape[0]
    adaptive = inner_tols is not None
    tols = tuple(inner_tols) if adaptive else (inner_tol,)
    u = np.zeros_like(rhs)
    rhs_norm = np.maximum(np.sqrt((rhs * rhs).sum(1)), 1e-30)
    best_res = np.full(B, np.inf)
    best_u = u.copy()
    total, last_k, prev, ti, cur_max, stall, stop = 0, 0, None, 0, int(max_inner), 0, 'max_outer'
    history = []
    for k in range(max_outer):
        r = rhs - apply_outer(u)
        res = np.sqrt((r * r).sum(1)) / rhs_norm
        last_k = k
        imp = res < best_res
        best_u = np.where(imp[:, None], u, best_u)
        best_res = np.where(imp, res, best_res)
        if res.max() < outer_tol:
            stop = 'converged'
            break
        tightened = False
        if adaptive:
            if prev is not None and res.max() > stagnation_factor * prev:
                if ti < len(tols) - 1:
                    ti += 1
                    cur_max = max(cur_max, int(adaptive_max_inner))
                    tightened = True
                    stall = 0
                elif res.max() > stall_factor * prev:
                    stall += 1
                else:
                    stall = 0
            else:
                stall = 0
            if stall >= stall_patience:
                stop = 'stalled'
                break
        du, n_in, inner_res = solve_pcg_batched(apply_inner, precond_inner, r, tol=tols[ti], max_iter=cur_max)
        total += int(n_in)
        history.append(dict(outer_iter=k, inner_tol=tols[ti], inner_max_iter=cur_max, inner_iters=int(n_in),
                            tightened=tightened, stall_count=stall))
        u = project(u + du)
        prev = float(res.max())
    if best_res.max() < outer_tol:
        stop = 'converged'
    return best_u, dict(outer_iters=last_k + 1, inner_iters=total, rel_res=best_res, stop_reason=stop,
                        history=history)


def backend_solve(apply, apply_precond, rhs, n, x0=None, rtol=1e-8, maxiter=4000):
    """Certified generic solve with restarts (``backend.py:123-144``)."""
    rhs = np.atleast_2d(np.asarray(rhs, float))
    norm = float(np.linalg.norm(rhs))
    atol = 1e-13 / math.sqrt(n)
    target = max(atol, rtol * norm)
    u = np.zeros_like(rhs) if x0 is None else hex8.project(x0)
    absolute = float(np.linalg.norm(rhs - apply(u)))
    iterations = restarts = 0
    if absolute > target:
        while iterations < maxiter:
            u, it, _ = solve_pcg_batched(apply, apply_precond, rhs, x0=u, tol=target / max(norm, 1e-30),
                                         max_iter=maxiter - iterations, update_mode='compiled',
                                         convergence_check_stride=4)
            iterations += it
            u = hex8.project(u)
            absolute = float(np.linalg.norm(rhs - apply(u)))
            if absolute <= target or not np.isfinite(absolute) or restarts >= 2:
                break
            restarts += 1
    info = dict(iterations=iterations, restarts=restarts, rhs_norm=norm, absolute_residual=absolute,
                relative_residual=absolute / max(norm, 1e-300), threshold=target,
                converged=bool(np.isfinite(absolute) and absolute <= target))
    return u, info
