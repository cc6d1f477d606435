"""General-C21 hybrid kernels and HybridFIOperator vs the oracle and closed forms (GPU)."""
import numpy as np
import pytest
import torch

import pgj
from pgj import materials as mat
from pgj.hybrid import HybridFIOperator, enable_fi_reference
from oracle import hex8, hybrid

pytestmark = pytest.mark.gpu
LAM, MU = 15 / 26, 5 / 13
C0 = hex8.isotropic(LAM, MU)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def spd(rng, scale=1.0):
    A = rng.standard_normal((6, 6))
    return scale * (A @ A.T + 6 * np.eye(6))


def random_rotation(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def multiphase(grid, seed, voxel=None):
    """Three phases (isotropic, local-frame orthotropic with per-voxel frames, anisotropic)."""
    rng = np.random.default_rng(seed)
    phase = rng.choice([0, 1, 2], size=grid).astype(np.int32)
    materials = {0: dict(kind='isotropic', E=1.0, nu=0.3),
                 1: dict(kind='orthotropic', frame='local', E1=10, E2=2, E3=1, nu12=0.3, nu13=0.2, nu23=0.25,
                         G12=1.5, G13=0.8, G23=0.6),
                 2: dict(kind='anisotropic', C=spd(rng, 0.5))}
    R = np.stack([random_rotation(rng) for _ in range(phase.size)]).reshape(*grid, 3, 3)
    pool, index, _ = mat.build_hybrid_material(phase, materials, orientation=R, device='cuda')
    vox = hex8.voxel_of(grid) if voxel is None else voxel
    return rng, pool, index, vox


CASES = [((6, 6, 6), None), ((3, 4, 5), (0.7, 1.1, 1.4)), ((8, 6, 10), None), ((5, 7, 4), None),
         ((2, 2, 2), None), ((20, 5, 17), None), ((9, 18, 33), None), ((3, 40, 2), None)]


@pytest.mark.parametrize('dense', [False, True], ids=['indirect', 'dense'])
@pytest.mark.parametrize('grid,voxel', CASES, ids=[str(g) for g, _ in CASES])
def test_apply_rhs_stress_match_oracle(grid, voxel, dense):
    rng, pool, index, vox = multiphase(grid, 11, voxel)
    if dense:
        pool = pool[index[0].long()].contiguous()
        index_h = None
        op = HybridFIOperator(pool, torch.empty(0, dtype=torch.int32, device='cuda'), grid, vox, dense=True)
    else:
        index_h = index[0].cpu().numpy()
        op = HybridFIOperator(pool, index, grid, vox)
    pool_h = pool.cpu().numpy()
    n = int(np.prod(grid))
    u = rng.normal(size=(3, 3 * n))
    ut = torch.tensor(u, device='cuda')
    ref = hybrid.apply(pool_h, index_h, u, grid, vox)
    batch = op.apply_batched(ut).cpu().numpy()
    assert rel(batch, ref) < 1e-13
    assert rel(op.apply_gather(ut).cpu().numpy(), ref) < 1e-13
    one = op.apply(ut[1:2].contiguous()).cpu().numpy()
    assert np.array_equal(one[0], batch[1])
    masked = op.apply_batched(ut, torch.tensor([1, 0, 1], device='cuda')).cpu().numpy()
    assert (masked[1] == 0).all() and np.array_equal(masked[[0, 2]], batch[[0, 2]])
    for m in [np.eye(6)[3], rng.normal(size=6)]:
        assert rel(op.compute_rhs(m).cpu().numpy(), hybrid.rhs(pool_h, index_h, m, grid, vox)) < 1e-13
        s = op.compute_avg_stress(ut[:1].contiguous(), m)[0]
        assert rel(s, hybrid.avg_stress(pool_h, index_h, u[:1], m, grid, vox)) < 1e-13
    p = op.project_zero_mean(ut[:1]).cpu().numpy()
    assert rel(p, hex8.project(u[:1])) < 1e-13


@pytest.mark.parametrize('grid', [(64, 64, 64), (48, 32, 40), (17, 33, 16)], ids=str)
def test_tiled_apply_equals_gather_at_size(grid):
    """Tiled and gather kernels at full size, with z-chunking and ragged tiles, several loads and a mask."""
    rng, pool, index, vox = multiphase(grid, 17)
    op = HybridFIOperator(pool, index, grid, vox)
    u = torch.tensor(rng.normal(size=(6, 3 * int(np.prod(grid)))), device='cuda')
    for L in (6, 1):
        a = op.apply_batched(u[:L].contiguous()).cpu().numpy()
        b = op.apply_gather(u[:L].contiguous()).cpu().numpy()
        assert rel(a, b) < 1e-13
    mask = torch.tensor([1, 1, 0, 1, 0, 1], device='cuda')
    m = op.apply_batched(u, mask).cpu().numpy()
    assert (m[[2, 4]] == 0).all() and np.array_equal(m[[0, 1, 3, 5]], op.apply_batched(u).cpu().numpy()[[0, 1, 3, 5]])


@pytest.mark.parametrize('n', [4, 32, 64])
def test_scalar_field_through_c21_equals_scalar_kernel(n):
    """a(x) C0 stored as dense C21 rows gives the scalar kernel's K, rhs and stress."""
    rng = np.random.default_rng(n)
    a = np.exp(rng.uniform(-4, 0, (n, n, n)))
    at = torch.tensor(a, device='cuda')
    op = pgj.Operator(at, None, LAM, MU)
    rows = at.reshape(-1, 1) * torch.tensor(hybrid.pack(C0), device='cuda')[None]
    hop = HybridFIOperator(rows.contiguous(), torch.empty(0, dtype=torch.int32, device='cuda'), (n, n, n),
                           op._voxel, dense=True)
    u = torch.tensor(rng.normal(size=(2, 3 * n ** 3)), device='cuda')
    assert rel(hop.apply_batched(u).cpu().numpy(), op.apply_batched(u).cpu().numpy()) < 1e-13
    m = rng.normal(size=6)
    assert rel(hop.compute_rhs(m).cpu().numpy(), op.compute_rhs(m).cpu().numpy()) < 1e-13
    one = u[:1].contiguous()
    assert rel(hop.compute_avg_stress(one, m), op.compute_avg_stress(one, m)) < 1e-13


def test_laminate_effective_tensor_end_to_end():
    """Rotated orthotropic / anisotropic / isotropic z-layers: the average-stress C^H of the solved
    cell problem equals the laminate closed form (the Q1 space holds the exact solution)."""
    rng = np.random.default_rng(21)
    grid = (16, 8, 8)
    layers = np.array([0] * 6 + [1] * 4 + [2] * 6, np.int32)
    phase = np.broadcast_to(layers[:, None, None], grid).copy()
    materials = {0: dict(kind='isotropic', E=1.0, nu=0.3),
                 1: dict(kind='orthotropic', frame='local', E1=10, E2=2, E3=1, nu12=0.3, nu13=0.2, nu23=0.25,
                         G12=1.5, G13=0.8, G23=0.6),
                 2: dict(kind='anisotropic', C=spd(rng, 0.3))}
    R = random_rotation(rng)
    pool, index, meta = mat.build_hybrid_material(phase, materials, orientation=R, device='cuda')
    assert meta['n_local_voxels'] == 4 * 64
    op = enable_fi_reference(HybridFIOperator(pool, index, grid, hex8.voxel_of(grid)))
    C, info = pgj.compute_effective_stiffness_batched(op, tol=1e-11, max_iter=3000)
    pool_h = pool.cpu().numpy()
    Cs = [hybrid.unpack(pool_h[0]), hybrid.unpack(pool_h[3]), hybrid.unpack(pool_h[2])]
    want = hybrid.laminate(Cs, [6, 4, 6])
    assert all(info[0][J]['converged'] for J in range(6))
    assert rel(C[0], want) < 1e-9


def test_reference_medium_and_preconditioned_solve():
    grid = (8, 8, 8)
    rng, pool, index, vox = multiphase(grid, 5)
    op = HybridFIOperator(pool, index, grid, vox)
    ref = op._compute_reference_medium('zeman')[0]
    C = hybrid.unpack(pool.cpu().numpy())[index[0].cpu().numpy()]
    mus = (C[:, 3, 3] + C[:, 4, 4] + C[:, 5, 5]) / 4
    Ks = (C[:, 0, 0] + C[:, 1, 1] + C[:, 2, 2] + 2 * (C[:, 0, 1] + C[:, 0, 2] + C[:, 1, 2])) / 9
    mu, K = (mus.max() + mus.min()) / 2, (Ks.max() + Ks.min()) / 2   # every voxel is sampled at n <= 16
    assert np.isclose(ref[3, 3], mu, rtol=1e-14) and np.isclose(ref[0, 1], max(0.0, K - 2 * mu / 3), rtol=1e-14)
    enable_fi_reference(op)
    rhs = op.compute_rhs(np.eye(6)[4])
    u, iters, res = pgj.solve_pcg_batched(op, rhs, tol=1e-10, max_iter=2000)
    f = rhs.cpu().numpy()
    r = f - hybrid.apply(pool.cpu().numpy(), index[0].cpu().numpy(), u.cpu().numpy(), grid, vox)
    assert np.linalg.norm(r) / np.linalg.norm(f) < 1e-9 and float(res.max()) < 1e-10


def test_argument_errors():
    grid = (4, 4, 4)
    _, pool, index, vox = multiphase(grid, 3)
    bad = index.clone()
    bad[0, 5] = pool.shape[0]
    with pytest.raises(ValueError, match='row map'):
        HybridFIOperator(pool, bad, grid, vox)
    with pytest.raises(ValueError, match='CUDA FP64'):
        HybridFIOperator(pool.float(), index, grid, vox)
    with pytest.raises(ValueError, match='dense storage'):
        HybridFIOperator(pool, index, grid, vox, dense=True)
    op = HybridFIOperator(pool, index, grid, vox)
    with pytest.raises(ValueError, match=r'\(1,3N\)'):
        op.apply(torch.zeros(2, 3 * 64, dtype=torch.float64, device='cuda'))
    with pytest.raises(ValueError, match='Six'):
        op.compute_rhs(np.zeros(5))


@pytest.mark.parametrize('n', [16, 24, 64, 128])
def test_fi_reference_mixed_precision_applies_in_fp32(n):
    """precision='mixed' (the reference's default) applies the FI Green operator in FP32 on compile-time cubes:
    within FP32 rounding of the oracle's FP64 Green (and not within FP64 rounding), returned in FP64; 'fp64'
    matches the oracle to 1e-13. Tolerance of the FP32 apply: 2e-5 of max|z| (FP32 transforms of 2n-point
    rows)."""
    from oracle import green
    grid = (n, n, n)
    rng = np.random.default_rng(17)
    phase = rng.choice([0, 1, 2], size=grid).astype(np.int32)
    materials = {0: dict(kind='isotropic', E=1.0, nu=0.3), 1: dict(kind='isotropic', E=10.0, nu=0.2),
                 2: dict(kind='anisotropic', C=spd(rng, 0.5))}
    pool, index, _ = mat.build_hybrid_material(phase, materials, device='cuda')
    vox = hex8.voxel_of(grid)
    mixed = enable_fi_reference(HybridFIOperator(pool, index, grid, vox))
    full = enable_fi_reference(HybridFIOperator(pool, index, grid, vox), precision='fp64')
    assert mixed._plan32 is not None and full._plan32 is None
    ref_C0 = np.asarray(mixed._formula_reference_C0)
    r = np.random.default_rng(n).standard_normal((1, 3 * n ** 3))
    rt = torch.tensor(r, device='cuda')
    zm, zf = mixed.apply_precond(rt), full.apply_precond(rt)
    assert zm.dtype == torch.float64 and zm.shape == rt.shape
    want = green.precond(r, green.symbol(grid, vox, ref_C0[0, 1], ref_C0[3, 3]), grid, None)
    assert rel(zf.cpu().numpy(), want) < 1e-13
    err = rel(zm.cpu().numpy(), want)
    assert 1e-10 < err < 2e-5, err
    # the FP32 result depends only on the FP32-rounded residual
    again = mixed.apply_precond(rt.float().double())
    assert torch.equal(again, zm)


def test_fi_reference_mixed_solve_and_other_grids():
    """The generic PCG with the FP32 Green apply still certifies in FP64 (the residual is FP64); a grid outside
    the compile-time cubes takes the FP64 apply for precision='mixed'."""
    grid = (16, 16, 16)
    _, pool, index, vox = multiphase(grid, 19)
    op = enable_fi_reference(HybridFIOperator(pool, index, grid, vox))
    rhs = op.compute_rhs(np.eye(6)[3])
    u, iters, res = pgj.solve_pcg_batched(op, rhs, tol=1e-10, max_iter=2000)
    f = rhs.cpu().numpy()
    r = f - hybrid.apply(pool.cpu().numpy(), index[0].cpu().numpy(), u.cpu().numpy(), grid, vox)
    assert np.linalg.norm(r) / np.linalg.norm(f) < 1e-9 and float(res.max()) < 1e-10
    grid = (6, 8, 10)
    _, pool, index, vox = multiphase(grid, 23)
    mixed = enable_fi_reference(HybridFIOperator(pool, index, grid, vox))
    full = enable_fi_reference(HybridFIOperator(pool, index, grid, vox), precision='fp64')
    assert mixed._plan32 is None
    rt = torch.tensor(np.random.default_rng(3).standard_normal((1, 3 * 480)), device='cuda')
    assert torch.equal(mixed.apply_precond(rt), full.apply_precond(rt))
