"""Pins the oracle's PCG recurrences against dense solves and forced failure paths (CPU)."""
import numpy as np
import pytest

from oracle import design, green, hex8, pcg

LAM, MU = 15 / 26, 5 / 13
C0 = hex8.isotropic(LAM, MU)


def dense_solutions(a, loads):
    """u = pinned-node solve then zero-mean projection (``exp/asg_scenarios_20261105/verify.py:19-27``)."""
    K = hex8.dense_matrix(a, C0)
    out = []
    for m in loads:
        f = hex8.rhs(a, m, C0)[0]
        u = np.zeros_like(f)
        u[3:] = np.linalg.solve(K[3:, 3:], f[3:])
        out.append(hex8.project(u[None])[0])
    return np.array(out)


@pytest.mark.parametrize('batch', [1, 3, 6])
def test_resident_solve_matches_dense(batch):
    n = 4
    rng = np.random.default_rng(51)
    a = design.simp(rng.uniform(.05, 1, (n, n, n)), 1e-4)
    s = green.scale(a, .5, C0)
    solver = pcg.ResidentPCG(a, C0, np.eye(6), batch=batch, s=s)
    infos, checks = solver.solve(rtol=1e-11)
    ref = dense_solutions(a, np.eye(6))
    err = np.linalg.norm(solver.solutions - ref) / np.linalg.norm(ref)
    assert err < 1e-8
    assert all(i['converged'] and i['relative_residual'] <= 1e-11 for i in infos)
    assert checks >= 1


def test_serial_and_parallel_agree():
    n = 6
    a = design.simp(design.filtered_noise(n, 20261104, 1), 1e-4)
    s = green.scale(a, .5, C0)
    runs = []
    for batch in (1, 3):
        solver = pcg.ResidentPCG(a, C0, np.eye(6), batch=batch, s=s)
        infos, _ = solver.solve(rtol=1e-10)
        runs.append((solver.solutions, [i['iterations'] for i in infos]))
    assert np.abs(runs[0][0] - runs[1][0]).max() < 1e-13 * np.abs(runs[0][0]).max()
    assert runs[0][1] == runs[1][1]


def test_zero_rhs_takes_zero_iterations():
    n = 4
    solver = pcg.ResidentPCG(np.full((n, n, n), .13), C0, np.eye(6), batch=3)
    infos, checks = solver.solve()
    assert all(i['iterations'] == 0 and i['converged'] for i in infos)
    assert checks == 2


def test_iteration_cap_raises_after_restarts(monkeypatch):
    n = 6
    a = design.simp(design.filtered_noise(n, 7, 1), 1e-6)
    monkeypatch.setattr(pcg, 'CAP', 3)
    solver = pcg.ResidentPCG(a, C0, np.eye(6)[:3], batch=3, s=None)
    with pytest.raises(RuntimeError) as info:
        solver.solve(rtol=1e-10)
    details = info.value.args[0][1]
    assert all(d['iterations'] == 3 for d in details)
    assert not any(d['converged'] for d in details)


def test_nan_latches_bad():
    n = 4
    a = design.simp(design.filtered_noise(n, 3, 1), 1e-4)
    solver = pcg.ResidentPCG(a, C0, np.eye(6)[3:4], batch=1)
    orig = solver.K
    calls = []

    def poisoned(p, active):
        q = orig(p, active)
        calls.append(1)
        if len(calls) == 2:
            q[:] = np.nan
        return q
    solver.K = poisoned
    with pytest.raises(RuntimeError) as info:
        solver.solve()
    assert info.value.args[0][1][0]['iterations'] == 2


def test_generic_pcg_and_backend_solve():
    n = 4
    rng = np.random.default_rng(731)
    a = np.exp(rng.uniform(-4, 0, (n, n, n)))
    G6 = green.symbol(a.shape, hex8.voxel_of(a.shape), LAM, MU)
    s = green.scale(a, 1.0, C0)
    apply = lambda u: hex8.apply(a, u, C0)
    pre = lambda r: green.precond(r, G6, a.shape, s)
    ref = dense_solutions(a, np.eye(6)[3:])
    for j, m in enumerate(np.eye(6)[3:]):
        f = hex8.rhs(a, m, C0)
        u, info = pcg.backend_solve(apply, pre, f, n, rtol=1e-12)
        assert info['converged']
        assert np.linalg.norm(u[0] - ref[j]) / np.linalg.norm(ref[j]) < 1e-9
        u2, it, res = pcg.solve_pcg_batched(apply, pre, f, tol=1e-10, max_iter=500)
        assert res[0] < 1e-10 and it < 500
        u3, it3, res3 = pcg.solve_pcg_batched(apply, pre, f, tol=1e-10, max_iter=500, update_mode='compiled',
                                              convergence_check_stride=4)
        assert it3 == it and np.array_equal(u3, u2)
    with pytest.raises(ValueError):
        pcg.solve_pcg_batched(apply, pre, f, update_mode='bogus')
    with pytest.raises(ValueError):
        pcg.solve_pcg_batched(apply, pre, f, convergence_check_stride=0)


def _spd_system(seed, n=40, cond=50.0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    K = q @ np.diag(np.geomspace(1.0, cond, n)) @ q.T
    return rng, K


def test_ir_reaches_the_dense_solution():
    rng, K = _spd_system(1)
    f = rng.standard_normal((2, K.shape[0]))
    u, info = pcg.solve_ir_batched(lambda x: x @ K, lambda x: x @ K, lambda r: r, lambda v: v, f,
                                   outer_tol=1e-10, inner_tol=0.1, max_inner=200)
    np.testing.assert_allclose(u, np.linalg.solve(K, f.T).T, rtol=1e-8, atol=1e-9)
    assert info['stop_reason'] == 'converged' and info['rel_res'].max() < 1e-10
    # an exact inner solve converges in one correction
    _, one = pcg.solve_ir_batched(lambda x: x @ K, lambda x: x @ K, lambda r: r, lambda v: v, f,
                                  outer_tol=1e-8, inner_tol=1e-13, max_inner=400)
    assert one['outer_iters'] == 2 and one['inner_iters'] > 0


@pytest.mark.parametrize('c,stop', [(2.0, 'converged'), (20.0, 'stalled')])
def test_adaptive_ir_schedule_of_a_scaled_inner_operator(c, stop):
    """Inner operator c K solved tightly: each correction is K^-1 r / c, so the outer residual
    contracts by exactly 1 - 1/c. c = 2 (ratio 0.5 > 0.3) tightens at every step until the last
    tolerance and still converges; c = 20 (ratio 0.95 > 0.9) stalls three steps after the last
    tightening."""
    rng, K = _spd_system(2)
    f = rng.standard_normal((1, K.shape[0]))
    u, info = pcg.solve_ir_batched(lambda x: x @ K, lambda x: c * (x @ K), lambda r: r, lambda v: v, f,
                                   outer_tol=1e-8, inner_tols=[1e-12, 1e-12, 1e-12], max_inner=400,
                                   adaptive_max_inner=400, max_outer=60)
    h = info['history']
    assert [e['tightened'] for e in h[:3]] == [False, True, True]
    assert info['stop_reason'] == stop
    if stop == 'stalled':
        assert [e['stall_count'] for e in h[2:]] == [0, 1, 2] and info['outer_iters'] == len(h) + 1
    else:
        assert info['outer_iters'] == int(np.ceil(np.log(1e-8) / np.log(0.5))) + 1
