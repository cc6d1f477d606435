"""HIP operator kernels vs the oracle: apply, rhs, energies, C^H, stress, diagonal, scale (GPU)."""
import numpy as np
import pytest
import torch

import pgj
from pgj import _lib
from oracle import design, green, hex8

pytestmark = pytest.mark.gpu
LAM, MU = 15 / 26, 5 / 13
C0 = hex8.isotropic(LAM, MU)

GRIDS = [((4, 4, 4), None), ((6, 6, 6), None), ((8, 8, 8), None), ((16, 16, 16), None),
         ((3, 4, 5), (0.7, 1.1, 1.4)), ((8, 6, 10), None), ((4, 6, 96), None), ((4, 4, 128), None)]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def make(grid, voxel, seed=731):
    rng = np.random.default_rng(seed)
    a = np.exp(rng.uniform(-4, 0, grid))
    op = pgj.Operator(torch.tensor(a, device='cuda'), voxel, LAM, MU)
    return rng, a, op, (hex8.voxel_of(grid) if voxel is None else voxel)


@pytest.mark.parametrize('grid,voxel', GRIDS, ids=[str(g) for g, _ in GRIDS])
def test_apply_matches_oracle(grid, voxel):
    rng, a, op, vox = make(grid, voxel)
    u = rng.normal(size=(3, 3 * a.size))
    ref = hex8.apply(a, u, C0, vox)
    ut = torch.tensor(u, device='cuda')
    fast = op.apply_batched(ut).cpu().numpy()
    slow = op.apply_reference(ut).cpu().numpy()
    assert rel(fast, ref) < 1e-14
    assert rel(slow, ref) < 1e-14
    one = op.apply(ut[1:2].contiguous()).cpu().numpy()
    assert np.array_equal(one[0], fast[1])   # a load alone and in the batch agree bitwise
    masked = op.apply_batched(ut, torch.tensor([True, False, True], device='cuda')).cpu().numpy()
    assert (masked[1] == 0).all() and np.array_equal(masked[0], fast[0]) and np.array_equal(masked[2], fast[2])


@pytest.mark.parametrize('grid,voxel', GRIDS[:5], ids=[str(g) for g, _ in GRIDS[:5]])
def test_rhs_energy_stress_diagonal(grid, voxel):
    rng, a, op, vox = make(grid, voxel, 5)
    for m in np.eye(6):
        assert rel(op.compute_rhs(m).cpu().numpy(), hex8.rhs(a, m, C0, vox)) < 1e-14
    u = rng.normal(size=(1, 3 * a.size)) * .05
    res = pgj.Resident.__new__(pgj.Resident)   # use the energy binding without a full solver
    res.op, res.loads = op, np.eye(6)
    sol = torch.tensor(np.repeat(u, 6, axis=0), device='cuda')
    e = res.energies(sol).cpu().numpy()
    assert rel(e, hex8.energies(np.repeat(u, 6, axis=0), grid, np.eye(6), C0, vox)) < 1e-13
    for m in np.eye(6):
        s = op.compute_avg_stress(torch.tensor(u, device='cuda'), m)[0]
        assert rel(s, hex8.avg_stress(a, u, m, C0, vox)) < 1e-13
    d = pgj.diagonal(op).cpu().numpy()
    assert rel(d, hex8.diagonal(a, C0, vox)) < 1e-14
    for alpha in (.5, 1.):
        s = pgj.scale(op, alpha).cpu().numpy()
        assert rel(s, green.scale(a, alpha, C0, vox)) < 1e-13
    pu = op.project_zero_mean(torch.tensor(u, device='cuda')).cpu().numpy()
    assert rel(pu, hex8.project(u)) < 1e-14


def test_effective_tensor_matches_oracle():
    grid = (8, 8, 8)
    rng, a, op, vox = make(grid, None, 11)
    u6 = rng.normal(size=(6, 3 * a.size)) * .01
    res = pgj.Resident.__new__(pgj.Resident)
    res.op, res.loads = op, np.eye(6)
    res.solutions = torch.tensor(u6, device='cuda')
    res.ch_ws = torch.empty(_lib.lib().pgj_effective_tensor_ws_bytes(op.grid) // 8 + 1, dtype=torch.float64,
                            device='cuda')
    res.ch_out = torch.empty(36, dtype=torch.float64, device='cuda')
    H = res.effective_tensor().cpu().numpy()
    ref = design.effective_tensor(a, u6, C0)
    assert rel(H, ref) < 1e-12


def test_operator_properties_at_64():
    """Size-independent checks at the 64^3 workload: symmetry and translation invariance."""
    n = 64
    a = torch.tensor(design.simp(design.filtered_noise(n, 20261104, 4), 1e-6), device='cuda')
    op = pgj.Operator(a)
    g = torch.Generator(device='cuda').manual_seed(731)
    u = torch.randn((2, 3 * n ** 3), generator=g, device='cuda', dtype=torch.float64)
    Ku = op.apply_batched(u)
    sym = float(((u[0] * Ku[1]).sum() - (u[1] * Ku[0]).sum()).abs() / ((u[0] * Ku[0]).sum() * (u[1] * Ku[1]).sum()).sqrt())
    assert sym < 1e-13
    shifted = (u.reshape(2, -1, 3) + torch.tensor([1., -2., .5], device='cuda', dtype=torch.float64)).reshape(2, -1)
    tr = float((op.apply_batched(shifted.contiguous()) - Ku).norm() / Ku.norm())
    assert tr < 1e-13
    ref = hex8.apply(a.cpu().numpy(), u[:1].cpu().numpy(), C0)
    assert rel(Ku[:1].cpu().numpy(), ref) < 1e-14


@pytest.mark.parametrize('grid,voxel', GRIDS, ids=[str(g) for g, _ in GRIDS])
@pytest.mark.parametrize('zc,w', [(1, 2), (2, 4), (4, 1)])
@pytest.mark.parametrize('plane', [0, 1], ids=['node', 'plane'])
def test_row_march_variants_match_oracle(grid, voxel, zc, w, plane):
    """Both row-march variants (node carry; plane carry with shared in-plane butterflies) against the oracle,
    several shapes."""
    rng, a, op, vox = make(grid, voxel, 99)
    u = rng.normal(size=(3, 3 * a.size))
    ref = hex8.apply(a, u, C0, vox)
    lib = _lib.lib()
    _lib.check(lib.pgj_hex8_set_shape(op.grid, 3, 8, 1, w, min(zc, grid[0]), plane))
    try:
        shape = (_lib.c_int * 5)()
        _lib.check(lib.pgj_hex8_get_shape(op.grid, 3, 8, shape))
        assert shape[3] == plane
        ut = torch.tensor(u, device='cuda')
        got = op.apply_batched(ut).cpu().numpy()
        assert rel(got, ref) < 1e-14
        masked = op.apply_batched(ut, torch.tensor([False, True, False], device='cuda')).cpu().numpy()
        assert (masked[0] == 0).all() and np.array_equal(masked[1], got[1])
    finally:
        _lib.check(lib.pgj_hex8_set_shape(op.grid, 3, 8, 0, 0, 0, 0))


@pytest.mark.parametrize('grid,voxel', GRIDS, ids=[str(g) for g, _ in GRIDS])
@pytest.mark.parametrize('zc,w', [(1, 2), (2, 4), (4, 1)])
def test_two_row_node_march_matches_oracle(grid, voxel, zc, w):
    """The node march with two element rows per wave (R = 2, a tuner-settable shape; odd ny and two x per lane,
    nx > 64, fall back to one row) against the oracle, with the load mask."""
    rng, a, op, vox = make(grid, voxel, 98)
    u = rng.normal(size=(3, 3 * a.size))
    ref = hex8.apply(a, u, C0, vox)
    lib = _lib.lib()
    _lib.check(lib.pgj_hex8_set_shape(op.grid, 3, 8, 2, w, min(zc, grid[0]), 0))
    try:
        shape = (_lib.c_int * 5)()
        _lib.check(lib.pgj_hex8_get_shape(op.grid, 3, 8, shape))
        assert shape[0] == (2 if grid[1] % 2 == 0 and grid[2] <= 64 else 1) and shape[3] == 0
        ut = torch.tensor(u, device='cuda')
        got = op.apply_batched(ut).cpu().numpy()
        assert rel(got, ref) < 1e-14
        masked = op.apply_batched(ut, torch.tensor([True, False, True], device='cuda')).cpu().numpy()
        assert (masked[1] == 0).all() and np.array_equal(masked[2], got[2])
    finally:
        _lib.check(lib.pgj_hex8_set_shape(op.grid, 3, 8, 0, 0, 0, 0))
