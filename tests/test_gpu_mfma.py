"""FP32 HEX8 apply as a dense Ke contraction on the matrix cores vs the oracle and the modal row-march (GPU).

Tolerance: relative to the largest output entry, 5e-6 against the FP64 oracle on the FP32-rounded inputs (FP32
accumulation of 24-term element products and 8-term node sums) and against the FP32 row-march (same inputs,
different summation order)."""
import numpy as np
import pytest
import torch

import pgj
from pgj import _lib
from oracle import hex8

pytestmark = pytest.mark.gpu
LAM, MU = 15 / 26, 5 / 13
C0 = hex8.isotropic(LAM, MU)
TOL = 5e-6

GRIDS = [((4, 4, 64), None), ((3, 8, 64), (0.7, 1.1, 1.4)), ((6, 12, 64), None), ((2, 4, 128), None),
         ((5, 4, 128), (1.0, 0.5, 2.0))]


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def to_soa(u, N):
    return u.reshape(u.shape[0], N, 3).transpose(1, 2).contiguous().reshape(u.shape[0], 3 * N)


@pytest.mark.parametrize('grid,voxel', GRIDS, ids=[str(g) for g, _ in GRIDS])
def test_mfma_apply_matches_oracle_and_rowmarch(grid, voxel):
    print(f'mfma apply {grid} voxel {voxel}', flush=True)
    rng = np.random.default_rng(731)
    a = np.exp(rng.uniform(-4, 0, grid)).astype(np.float32).astype(np.float64)
    u = rng.normal(size=(3, 3 * a.size)).astype(np.float32)
    op = pgj.Operator(torch.tensor(a, device='cuda'), voxel, LAM, MU, dtype=torch.float32)
    vox = hex8.voxel_of(grid) if voxel is None else voxel
    ref = hex8.apply(a, u.astype(np.float64), C0, vox)
    ut = torch.tensor(u, device='cuda')
    got = op.apply_batched_mfma(ut)
    assert rel(got.double().cpu().numpy(), ref) < TOL
    assert rel(got.double().cpu().numpy(), op.apply_batched(ut).double().cpu().numpy()) < TOL
    # component-major vectors give the same numbers bitwise (same LDS contents, same arithmetic)
    N = a.size
    soa = torch.empty_like(ut)
    _lib.call('pgj_hex8_apply32_mfma', op.grid, 3, _lib.ptr(op._material32), _lib.ptr(to_soa(ut, N)), None,
              _lib.ptr(soa), 1, _lib.stream())
    assert torch.equal(soa, to_soa(got, N))
    # an inactive load gives exact zeros; the others are unchanged
    masked = op.apply_batched_mfma(ut, torch.tensor([1, 0, 1], device='cuda'))
    assert (masked[1] == 0).all() and torch.equal(masked[0], got[0]) and torch.equal(masked[2], got[2])


def test_mfma_apply_at_64_cube_six_loads():
    """The bench workload's shape (64^3, six loads, contrast 1e-6): against the row-march, plus symmetry
    v.K u = u.K v and K applied to a rigid translation = 0, both to FP32 accuracy."""
    n = 64
    print('mfma apply 64^3 x 6', flush=True)
    from pgj.density import filtered_noise, simp
    a = simp(filtered_noise(n, 20261104, 4), 1e-6)
    op = pgj.Operator(torch.tensor(a, device='cuda'), dtype=torch.float32)
    g = torch.Generator(device='cuda').manual_seed(731)
    u = torch.randn(6, 3 * n ** 3, dtype=torch.float32, device='cuda', generator=g)
    got = op.apply_batched_mfma(u)
    rm = op.apply_batched(u)
    assert rel(got.double().cpu().numpy(), rm.double().cpu().numpy()) < TOL
    Ku0, Ku1 = got[0].double(), got[1].double()
    lhs, rhs = float(u[1].double() @ Ku0), float(u[0].double() @ Ku1)
    assert abs(lhs - rhs) <= 1e-5 * float(u[0].double().norm() * Ku0.norm())
    t = torch.zeros(1, 3 * n ** 3, dtype=torch.float32, device='cuda')
    t[0, 1::3] = 1.0
    assert float(op.apply_batched_mfma(t).abs().max()) <= 1e-6 * float(got.abs().max())


def test_mfma_apply_rejects_unsupported_grids():
    op = pgj.Operator(torch.ones((4, 6, 64), dtype=torch.float64, device='cuda'), dtype=torch.float32)
    with pytest.raises(ValueError, match='ny divisible by 4'):
        op.apply_batched_mfma(torch.zeros(1, 3 * op.mesh.n_nodes, dtype=torch.float32, device='cuda'))
    op = pgj.Operator(torch.ones((4, 4, 32), dtype=torch.float64, device='cuda'), dtype=torch.float32)
    with pytest.raises(ValueError, match='nx = 64 or 128'):
        op.apply_batched_mfma(torch.zeros(1, 3 * op.mesh.n_nodes, dtype=torch.float32, device='cuda'))
    op64 = pgj.Operator(torch.ones((4, 4, 64), dtype=torch.float64, device='cuda'))
    with pytest.raises(ValueError):
        op64.apply_batched_mfma(torch.zeros(1, 3 * op64.mesh.n_nodes, dtype=torch.float64, device='cuda'))
