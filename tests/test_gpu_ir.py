"""FP32 operator, FP32 generic PCG and mixed-precision iterative refinement vs the oracle (GPU)."""
import numpy as np
import pytest
import torch

import pgj
from pgj import ir
from pgj.density import filtered_noise, simp
from oracle import green, hex8, pcg as opcg

pytestmark = pytest.mark.gpu
LAM, MU = 15 / 26, 5 / 13
C0 = hex8.isotropic(LAM, MU)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def design(n, seed=5, emin=1e-3):
    return simp(filtered_noise(n, seed, 1), emin)


def ops(a, scale_inner=1.0):
    at = torch.tensor(a, device='cuda')
    op64 = pgj.build(at)
    op32 = pgj.build(at * scale_inner, dtype=torch.float32)
    return op64, op32


@pytest.mark.parametrize('n', [8, 16, 32])
def test_fp32_apply_matches_fp64(n):
    a = design(n)
    op64, op32 = ops(a)
    u = torch.tensor(np.random.default_rng(n).normal(size=(3, 3 * n ** 3)), device='cuda')
    ref = hex8.apply(a, u.cpu().numpy(), C0)
    assert rel(op64.apply_batched(u).cpu().numpy(), ref) < 1e-13
    got = op32.apply_batched(u.float()).double().cpu().numpy()
    assert rel(got, ref) < 3e-6
    one = op32.apply(u[1:2].float().contiguous())
    assert one.dtype == torch.float32 and torch.equal(one[0], op32.apply_batched(u.float())[1])
    z = op32.apply_precond(u[:1].float().contiguous())
    assert z.dtype == torch.float32
    assert rel(z.double().cpu().numpy(), op64.apply_precond(u[:1].contiguous()).cpu().numpy()) < 1e-6


def test_fp32_generic_pcg_converges():
    n = 16
    a = design(n)
    op64, op32 = ops(a)
    f = op64.compute_rhs(np.eye(6)[3])
    u32, it, res = pgj.solve_pcg_batched(op32, f.float(), tol=1e-4, max_iter=500)
    assert u32.dtype == torch.float32 and float(res.max()) < 1e-4
    r = f.cpu().numpy() - hex8.apply(a, u32.double().cpu().numpy(), C0)
    assert np.linalg.norm(r) / np.linalg.norm(f.cpu().numpy()) < 2e-4
    G6 = green.symbol(a.shape, hex8.voxel_of(a.shape), LAM, MU)
    _, it_ref, _ = opcg.solve_pcg_batched(lambda x: hex8.apply(a, x, C0), lambda x: green.precond(x, G6, a.shape),
                                          f.cpu().numpy(), tol=1e-4, max_iter=500)
    assert abs(it - it_ref) <= 2


def test_ir_fp32_inner_reaches_fp64_accuracy():
    n = 16
    a = design(n)
    op64, op32 = ops(a)
    f = op64.compute_rhs(np.eye(6)[5])
    u, info = ir.solve_ir_batched(op64, op32, f, outer_tol=1e-9, inner_tol=0.1, max_inner=200)
    assert info['converged'] and info['stop_reason'] == 'converged' and info['rel_res'].max() < 1e-9
    fh = f.cpu().numpy()
    r = fh - hex8.apply(a, u.cpu().numpy(), C0)
    assert np.linalg.norm(r) / np.linalg.norm(fh) < 1e-9
    G6 = green.symbol(a.shape, hex8.voxel_of(a.shape), LAM, MU)
    _, ref = opcg.solve_ir_batched(lambda x: hex8.apply(a, x, C0), lambda x: hex8.apply(a, x, C0),
                                   lambda x: green.precond(x, G6, a.shape), hex8.project, fh, outer_tol=1e-9,
                                   inner_tol=0.1, max_inner=200)
    assert abs(info['outer_iters'] - ref['outer_iters']) <= 1


@pytest.mark.parametrize('c,stop', [(2.0, 'converged'), (20.0, 'stalled')])
def test_adaptive_ir_schedule_matches_oracle(c, stop):
    """Inner operator c K (FP32): the outer residual contracts by 1 - 1/c per step, which fixes the
    tightening and stall schedule; it must equal the oracle's."""
    n = 8
    a = design(n)
    op64, op32 = ops(a, c)
    f = op64.compute_rhs(np.eye(6)[0])
    kw = dict(outer_tol=1e-8, inner_tols=(1e-5, 1e-5, 1e-5), max_inner=400, adaptive_max_inner=400, max_outer=60)
    u, info = ir.solve_ir_adaptive_batched(op64, op32, f, **kw)
    G6 = green.symbol(a.shape, hex8.voxel_of(a.shape), LAM, MU)
    _, ref = opcg.solve_ir_batched(lambda x: hex8.apply(a, x, C0), lambda x: c * hex8.apply(a, x, C0),
                                   lambda x: green.precond(x, G6, a.shape) / c, hex8.project, f.cpu().numpy(), **kw)
    assert info['stop_reason'] == ref['stop_reason'] == stop
    assert info['outer_iters'] == ref['outer_iters']
    assert [h['tightened'] for h in info['history']] == [h['tightened'] for h in ref['history']]
    assert [h['stall_count'] for h in info['history']] == [h['stall_count'] for h in ref['history']]


def test_ir_effective_tensor_matches_resident():
    n = 16
    a = design(n, 9)
    op64, op32 = ops(a)
    C, info = ir.compute_effective_stiffness_ir_batched(op64, op32, outer_tol=1e-10, inner_tol=0.1, max_inner=200)
    res = pgj.Resident(n, load_mode='parallel', batch=6).capture()
    res.material(torch.tensor(a, device='cuda'))
    res.select(0.0)
    res.solve(rtol=1e-10)
    H = res.effective_tensor().cpu().numpy()
    assert rel(C[0], H) < 1e-8
    assert all(e['rel_res'] < 1e-10 for e in info[0])
    Ca, infoa = ir.compute_effective_stiffness_ir_adaptive_batched(op64, op32, outer_tol=1e-10, max_inner=200)
    assert rel(Ca[0], H) < 1e-8 and all(e['converged'] for e in infoa[0])
